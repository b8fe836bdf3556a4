"""Torch restatements of the MaterialGAN caller's network layers, in any dtype,
from the weights of the project's modules: the checker of the matrix-core
convolution kernels (test infrastructure, not a product path).

* conv3x3_ref: F.conv2d(x, w, padding=1) + noise_strength * noise + bias,
  then ReLU or leaky ReLU(0.2) x sqrt(2) (ModulateConvBlock.forward,
  src/higan_models/stylegan2_generator_network.py:694-755).
* upsample2_ref: zero insertion and the [1, 3, 3, 1] filter with gain 4 as
  two depthwise convolutions (UpsamplingLayer, :419-479).
* synthesis_ref: the generator's synthesis pass (modulated weights, the
  blocks above, toRGB skip sums) with every parameter cast to the dtype.
* vgg_features_ref / vgg_shaped_ref: the two VGG stacks (src/descriptor.py:7-79).
"""
import math

import torch
import torch.nn.functional as F

_BLUR = torch.tensor([1.0, 3.0, 3.0, 1.0])


def act_ref(v, act, mask=None):
    """The activation; with `mask` (a boolean tensor like v) the branch of
    each element is taken from the mask instead of the sign of v."""
    if mask is None:
        if act == "relu":
            return torch.relu(v)
        if act == "lrelu":
            return F.leaky_relu(v, 0.2) * math.sqrt(2.0)
        return v
    if act == "relu":
        return v * mask.to(v.dtype)
    if act == "lrelu":
        return torch.where(mask, v, 0.2 * v) * math.sqrt(2.0)
    return v


def conv3x3_ref(x, w, bias=None, noise=None, noise_strength=0.0, act="none", mask=None, record=None):
    """With `record` (a list) the pre-activation is appended to it."""
    y = F.conv2d(x, w.to(x.dtype), padding=1)
    if noise is not None:
        y = y + noise_strength * noise.to(x.dtype)
    if bias is not None:
        y = y + bias.to(x.dtype).view(1, -1, 1, 1)
    if record is not None:
        record.append(y.detach())
    return act_ref(y, act, mask)


def upsample2_ref(x):
    n, c, h, w = x.shape
    up = x.new_zeros(n, c, 2 * h, 2 * w)
    up[:, :, ::2, ::2] = x
    k = (_BLUR / _BLUR.sum() * 2.0).to(x)                            # per axis, gain 2
    up = F.pad(up, (2, 1, 2, 1), mode="constant")
    up = F.conv2d(up, k.view(1, 1, 1, 4).repeat(c, 1, 1, 1), groups=c)
    return F.conv2d(up, k.view(1, 1, 4, 1).repeat(c, 1, 1, 1), groups=c)


def _dense(layer, x):
    return F.linear(x, layer.weight.to(x.dtype) * layer.scale, layer.bias.to(x.dtype) * layer.lr_mul)


def modulated_weight_ref(conv, w):
    style = _dense(conv.affine, w)
    wt = conv.weight.to(w.dtype) * conv.scale * style.view(1, -1, 1, 1)
    if conv.demodulate:
        wt = wt * torch.rsqrt((wt * wt).sum(dim=(1, 2, 3), keepdim=True) + 1e-8)
    return wt


def synthesis_ref(gen, wp, noises, masks=None, record=None):
    """StyleGAN2Generator.synthesis in the dtype of wp (the noise maps are
    cast to it); masks / record as conv3x3_ref, one per 3x3 layer."""
    dt = wp.dtype
    x = gen.const.to(dt)
    rgb = None
    wi = 0
    for i, conv in enumerate(gen.convs):
        wt = modulated_weight_ref(conv, wp[:, min(wi, gen.num_layers - 1)])
        if conv.upsample:
            x = upsample2_ref(x)
        x = conv3x3_ref(x, wt, gen.bias[i], noises[i].to(dt), gen.noise_strength_values[i], "lrelu",
                        None if masks is None else masks[i], record)
        wi += 1
        if i == 0 or i % 2 == 0:
            to_rgb = gen.to_rgb[(i + 1) // 2]
            t = F.conv2d(x, modulated_weight_ref(to_rgb, wp[:, min(wi, gen.num_layers - 1)]))
            rgb = t if rgb is None else upsample2_ref(rgb) + t
    return rgb


def gram_ref(f):
    """Gram matrices per image, F F^T / (c h w), by torch autograd of bmm
    (the transposed operand made contiguous first)."""
    n, c, h, w = f.shape
    f = f.reshape(n, c, h * w)
    return torch.bmm(f, f.transpose(1, 2).contiguous()) / (c * h * w)


def vgg_shaped_ref(net, x, masks=None, record=None):
    """compose.VGGShapedFeatures.forward in x's dtype; masks / record as
    conv3x3_ref, one per convolution."""
    from svbrdf_amd import compose
    feats = []
    for step in net.plan:
        if step == "P":
            x = F.avg_pool2d(x, 2)
        else:
            c = net.convs[step]
            x = conv3x3_ref(x, c.weight, c.bias, act="relu", mask=None if masks is None else masks[step],
                            record=record)
            if step in compose.STYLE_LAYERS:
                feats.append(x)
    return feats


def vgg_features_ref(vgg, x, masks=None, record=None):
    """materialgan.VGGLoss.compute_feature_vector(normalize(x)) in x's dtype;
    masks / record as conv3x3_ref, one per convolution."""
    x = (x - vgg.mean.to(x.dtype)) / vgg.std.to(x.dtype)
    feats = []
    for step in vgg.plan:
        if step == "P":
            x = F.avg_pool2d(x, 2)
            continue
        c = vgg.convs[step]
        x = conv3x3_ref(x, c.weight, c.bias, act="relu", mask=None if masks is None else masks[step], record=record)
        if step in vgg.HOOKS:
            feats.append(x.flatten() * vgg.weights[len(feats)])
    return torch.cat(feats)
