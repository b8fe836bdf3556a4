"""Render adjoint composed with a feature-space loss.

The reference's second caller of the renderer (MaterialGAN,
src/materialgan.py:136-152) renders, passes the images through a VGG19
feature extractor (src/descriptor.py:7-79: pretrained VGG19, max-pool
replaced by average-pool) and back-propagates a feature loss through the
renderer.  No pretrained weights are available offline, so this module
builds a VGG-19-SHAPED stack with seeded He-normal weights: conv1_1 ...
conv5_1 (3x3 convolutions, ReLU, 2x2 average pooling between stages).
The loss is the L2 distance of Gram matrices at relu1_1, relu2_1,
relu3_1, relu4_1, relu5_1 plus a pixel L2 term.  The renderer's adjoint
kernel receives dL/dimages from the feature stack's backward (the
convolutions' input gradients are the matrix-core kernel of
:mod:`svbrdf_amd.convnet`) through :class:`svbrdf_amd.RenderFunction`.
"""
import torch

from . import convnet

# (out_channels) per conv layer, 'P' = 2x2 average pool; up to conv5_1
VGG19_TO_CONV5_1 = [64, 64, "P", 128, 128, "P", 256, 256, 256, 256, "P", 512, 512, 512, 512, "P", 512]
STYLE_LAYERS = (0, 2, 4, 8, 12)   # indices of conv1_1, conv2_1, conv3_1, conv4_1, conv5_1 among the convs


class VGGShapedFeatures(torch.nn.Module):
    """Random-init (He-normal, seeded) VGG-19-shaped conv stack up to conv5_1;
    each 3x3 convolution with its ReLU is a :class:`convnet.Conv3x3` layer
    (the matrix-core kernel, forward and input gradient)."""

    def __init__(self, seed=4, device="cuda"):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        self.convs, self.plan = [], []
        cin = 3
        for item in VGG19_TO_CONV5_1:
            if item == "P":
                self.plan.append("P")
                continue
            std = (2.0 / (cin * 9)) ** 0.5
            wgt = torch.randn((item, cin, 3, 3), generator=g) * std
            self.convs.append(convnet.Conv3x3(wgt.to(device), torch.zeros(item, device=device), act="relu"))
            self.plan.append(len(self.convs) - 1)
            cin = item

    def forward(self, x):
        feats = []
        for step in self.plan:
            if step == "P":
                x = convnet.avgpool2(x)
            else:
                x = self.convs[step](x)
                if step in STYLE_LAYERS:
                    feats.append(x)
        return feats


class _Gram(torch.autograd.Function):
    """G = F F^T / (c h w) per image of f [n, c, h, w] (a c x c matrix per
    image), with dL/dF = (dG + dG^T) F / (c h w).  Per-image 2-D products
    read the features as they lie in memory: no transposed copy of the
    feature map forward, no strided accumulation of a transposed gradient
    backward."""

    @staticmethod
    def forward(ctx, f):
        n, c, h, w = f.shape
        fm = f.reshape(n, c, h * w)
        g = torch.stack([torch.mm(fm[i], fm[i].t()) for i in range(n)]) / (c * h * w)
        ctx.save_for_backward(f)
        return g

    @staticmethod
    def backward(ctx, dg):
        (f,) = ctx.saved_tensors
        n, c, h, w = f.shape
        s = (dg + dg.transpose(1, 2)).contiguous() / (c * h * w)
        return torch.bmm(s, f.reshape(n, c, h * w)).reshape(n, c, h, w)


def gram(f):
    return _Gram.apply(f)


class FeatureLoss:
    """L2 of Gram matrices on the 5 style layers + pixel L2 (weights w_gram,
    w_pixel); gram_fn computes the Gram matrices (default: gram)."""

    def __init__(self, net, targets, w_gram=1.0, w_pixel=1.0, gram_fn=None):
        self.net = net
        self.gram = gram if gram_fn is None else gram_fn
        self.targets = targets.detach()
        with torch.no_grad():
            self.target_grams = [self.gram(f) for f in net(self.targets)]
        self.w_gram, self.w_pixel = w_gram, w_pixel

    def __call__(self, images):
        feats = self.net(images)
        loss = self.w_pixel * torch.nn.functional.mse_loss(images, self.targets)
        for f, gt in zip(feats, self.target_grams):
            loss = loss + self.w_gram * torch.nn.functional.mse_loss(self.gram(f), gt)
        return loss
