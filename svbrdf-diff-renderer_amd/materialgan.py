"""The MaterialGAN caller of the renderer (src/materialgan.py:21-153), composed
with the HIP render adjoint.

MaterialGAN optimises a StyleGAN2 latent (W+, 14 x 512 at 256^2) and the
per-layer noise maps so that the 9-channel SVBRDF the generator synthesises
renders like the target photos: image MSE plus a VGG-19 feature loss
(src/descriptor.py:7-79), alternating "L" epochs (latent, feature weights
[1, 1, 4, 8]/14) and "N" epochs (noise, [8, 8, 2, 1]/19).

The pretrained generator and VGG-19 weights are network downloads
(src/materialgan.py:37-56, src/descriptor.py:13) and are not available
offline, so both networks here have the published SHAPES with seeded random
weights:

* :class:`StyleGAN2Generator` -- mapping (8 x 512 fully connected, leaky
  ReLU), W+ with 14 layers at 256^2, synthesis from a learned 4 x 4 constant
  through modulated / demodulated 3 x 3 convolutions with noise injection,
  x2 upsampling with the [1, 3, 3, 1] binomial filter per resolution, and the
  "skip" output (9-channel toRGB at every resolution, upsampled and summed);
  the architecture of Karras et al., "Analyzing and Improving the Image
  Quality of StyleGAN" (2020), written for batch 1 (the modulation folds into
  the convolution weights).
* :class:`VGGLoss` -- VGG-19 features with max pooling replaced by average
  pooling, the outputs of layers 1, 3, 13, 22 (relu1_1, relu1_2, relu3_2,
  relu4_2) flattened, weighted and compared by MSE after ImageNet
  normalisation (std [0.229, 0.224, 0.255] as the reference writes it).
* :class:`MaterialGANOptim` -- init_from, load_targets, compute_feature_loss,
  latent_to_textures (clamps of src/materialgan.py:91-100) and optim.

The 3x3 convolutions of both networks run on the matrix-core kernels of
:mod:`svbrdf_amd.convnet` (forward, input gradient and the generator's
weight gradient, which the latent's gradient needs), with the noise, bias
and leaky-ReLU epilogue fused; in the generator the style modulation and the
demodulation happen inside the HIP weight packing, and dL/dstyle comes from
the HIP modulation backward.  The x2 upsampling is a HIP kernel too.  The
renderer is :class:`svbrdf_amd.Microfacet` (forward and adjoint kernels).
The 1x1 toRGB convolutions (9 output channels: plain GEMMs, torch.matmul)
and the fully connected mapping layers stay torch operations.
"""
import datetime
import math
import pathlib

import numpy as np
import torch
import torch.nn.functional as F

from . import convnet
from .optim import SvbrdfOptim

RES_LAYERS = {256: 14}


def _lrelu(x):
    return F.leaky_relu(x, 0.2) * math.sqrt(2.0)


class _Dense(torch.nn.Module):
    """Fully connected layer with the weight scale applied at run time."""

    def __init__(self, din, dout, g, bias_init=0.0, lr_mul=1.0):
        super().__init__()
        self.weight = torch.nn.Parameter(torch.randn(dout, din, generator=g) / lr_mul)
        self.bias = torch.nn.Parameter(torch.full((dout,), float(bias_init)))
        self.scale = lr_mul / math.sqrt(din)
        self.lr_mul = lr_mul

    def forward(self, x):
        # x W^T scale + b lr_mul as one GEMM with alpha (no pass over the
        # weight to scale it); lr_mul is 1 in the style layers
        b = self.bias if self.lr_mul == 1.0 else self.bias * self.lr_mul
        if x.dim() == 2:
            return torch.addmm(b, x, self.weight.t(), alpha=self.scale)
        return F.linear(x, self.weight * self.scale, b)


class _ModConv(torch.nn.Module):
    """Modulated convolution for one sample: weights scaled by the style of
    w per input channel, demodulated per output channel (optional), optional
    x2 upsampling (zero insertion + [1,3,3,1] blur, gain 4) before a 3x3
    convolution; 3x3 layers take the block's noise / bias / activation
    epilogue (ModulateConvBlock.forward,
    src/higan_models/stylegan2_generator_network.py:694-755)."""

    def __init__(self, cin, cout, k, g, demodulate=True, upsample=False):
        super().__init__()
        self.weight = torch.nn.Parameter(torch.randn(cout, cin, k, k, generator=g))
        self.affine = _Dense(512, cin, g, bias_init=1.0)
        self.scale = 1.0 / math.sqrt(cin * k * k)
        self.demodulate, self.upsample, self.k = demodulate, upsample, k

    def modulated_weight(self, w):
        style = self.affine(w)                                       # [1, cin]
        wt = self.weight * (style * self.scale).view(1, -1, 1, 1)
        if self.demodulate:
            wt = wt * torch.rsqrt((wt * wt).sum(dim=(1, 2, 3), keepdim=True) + 1e-8)
        return wt

    def forward(self, x, w, bias=None, noise=None, noise_strength=0.0, act="none"):
        if self.k == 1:
            # a 1x1 convolution is a plain GEMM: [Cout, Cin] x [Cin, H W] per
            # sample (torch.mm for the one sample: no broadcast copies)
            wt = self.modulated_weight(w)
            n, c, h, wd = x.shape
            if n == 1:
                return torch.mm(wt.view(wt.shape[0], c), x.reshape(c, h * wd)).view(1, -1, h, wd)
            return torch.matmul(wt.view(wt.shape[0], c), x.reshape(n, c, h * wd)).view(n, -1, h, wd)
        # 3x3: the modulation and demodulation happen inside the HIP weight
        # packing, and the style's gradient comes from the HIP modulation
        # backward, instead of torch ops on the whole weight tensor
        style = self.affine(w)
        if self.upsample:
            x = convnet.upsample2(x)
        return convnet.modulated_conv3x3(x, style, self.weight, self.scale, self.demodulate, bias, noise,
                                         noise_strength, act)


class StyleGAN2Generator(torch.nn.Module):
    """StyleGAN2 ("skip") generator shape at 256^2 with 9 output channels;
    ``mapping``, ``truncation`` and ``synthesis`` as the reference's net
    attributes (src/materialgan.py:61-62,92)."""

    def __init__(self, resolution=256, channels=9, seed=7, device="cuda"):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        self.resolution, self.channels = resolution, channels
        self.num_layers = RES_LAYERS[resolution]
        nf = lambda r: min(32768 // r, 512)
        self.map_layers = torch.nn.ModuleList([_Dense(512, 512, g, lr_mul=0.01) for _ in range(8)])
        self.const = torch.nn.Parameter(torch.randn(1, nf(4), 4, 4, generator=g))
        self.convs = torch.nn.ModuleList([_ModConv(nf(4), nf(4), 3, g)])
        self.to_rgb = torch.nn.ModuleList([_ModConv(nf(4), channels, 1, g, demodulate=False)])
        self.noise_strength = torch.nn.ParameterList([torch.nn.Parameter(torch.tensor(0.1))])
        self.conv_res = [4]
        r = 8
        while r <= resolution:
            self.convs.append(_ModConv(nf(r // 2), nf(r), 3, g, upsample=True))
            self.convs.append(_ModConv(nf(r), nf(r), 3, g))
            self.to_rgb.append(_ModConv(nf(r), channels, 1, g, demodulate=False))
            self.noise_strength.append(torch.nn.Parameter(torch.tensor(0.1)))
            self.noise_strength.append(torch.nn.Parameter(torch.tensor(0.1)))
            self.conv_res += [r, r]
            r *= 2
        self.bias = torch.nn.ParameterList([torch.nn.Parameter(torch.zeros(c.weight.shape[0])) for c in self.convs])
        # frozen: read once here, not from the device at every layer of every call
        self.noise_strength_values = [float(p) for p in self.noise_strength]
        self.register_buffer("w_avg", torch.zeros(512))
        self.to(device)
        for p in self.parameters():
            p.requires_grad_(False)

    def mapping(self, z):
        x = z * torch.rsqrt((z * z).mean(dim=1, keepdim=True) + 1e-8)
        for layer in self.map_layers:
            x = _lrelu(layer(x))
        return x

    def truncation(self, w, psi=1.0):
        """W -> W+ ([1, num_layers, 512]); psi = 1 keeps w (as the reference's
        generator is configured, truncation_psi = 1.0)."""
        wp = w.view(-1, 1, 512).repeat(1, self.num_layers, 1)
        return self.w_avg.view(1, 1, 512) + psi * (wp - self.w_avg.view(1, 1, 512))

    def make_noises(self, gen=None):
        """One [1, 1, r, r] noise map per convolution layer."""
        return [torch.randn((1, 1, r, r), generator=gen).to(self.const.device) for r in self.conv_res]

    def synthesis(self, wp, noises):
        """W+ [1, num_layers, 512] and the per-layer noise maps -> [1, C, R, R]."""
        x = self.const
        rgb = None
        wi = 0
        # one unbind instead of an index per layer: its backward is one stack
        # of the per-layer gradients, not a zero-filled copy per layer
        ws = wp.unbind(1)
        for i, conv in enumerate(self.convs):
            # conv + noise_strength * noise + bias, leaky ReLU x sqrt(2), in one kernel
            x = conv(x, ws[min(wi, self.num_layers - 1)], bias=self.bias[i], noise=noises[i],
                     noise_strength=self.noise_strength_values[i], act="lrelu")
            wi += 1
            if i == 0 or i % 2 == 0:
                # toRGB after each resolution's last convolution, skip-summed
                t = self.to_rgb[(i + 1) // 2](x, ws[min(wi, self.num_layers - 1)])
                rgb = t if rgb is None else convnet.upsample2(rgb) + t
        return rgb


class VGGLoss(torch.nn.Module):
    """src/descriptor.py:7-79 with seeded random VGG-19 weights; the 3x3
    convolutions with their ReLU are :class:`convnet.Conv3x3` layers."""

    LAYERS = [64, 64, "P", 128, 128, "P", 256, 256, 256, 256, "P", 512, 512]   # conv1_1 .. conv4_2
    HOOKS = (0, 1, 5, 9)              # the convolutions whose ReLU outputs are relu1_1, relu1_2, relu3_2, relu4_2

    def __init__(self, device, weights=np.array([1, 1, 1, 1]) / 4, seed=4):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        self.plan, self.convs, cin = [], [], 3
        for item in self.LAYERS:
            if item == "P":
                self.plan.append("P")           # max pooling replaced by average pooling, descriptor.py:18-20
                continue
            wgt = torch.randn((item, cin, 3, 3), generator=g) * math.sqrt(2.0 / (cin * 9))
            self.convs.append(convnet.Conv3x3(wgt.to(device), torch.zeros(item, device=device), act="relu"))
            self.plan.append(len(self.convs) - 1)
            cin = item
        self.weights = [float(x) for x in weights]
        self.mean = torch.tensor([0.485, 0.456, 0.406], device=device).view(1, 3, 1, 1)
        self.std = torch.tensor([0.229, 0.224, 0.255], device=device).view(1, 3, 1, 1)

    def normalize(self, im):
        return (im - self.mean) / self.std

    def _features(self, x):
        feats = []
        for step in self.plan:
            if step == "P":
                x = convnet.avgpool2(x)
                continue
            x = self.convs[step](x)
            if step in self.HOOKS:
                feats.append(x.flatten())
        return feats

    def compute_feature_vector(self, x):
        return torch.cat([f * w for f, w in zip(self._features(x), self.weights)])

    def load(self, im):
        with torch.no_grad():
            feats = self._features(self.normalize(im))
            self.im_feature = torch.cat([f * w for f, w in zip(feats, self.weights)])
            self._targets = feats

    def forward(self, x):
        # mean((w_l f_l - w_l t_l)^2) over the concatenated vector, taken per
        # segment as w_l^2 sum((f_l - t_l)^2): no weighted copies of the
        # feature maps and no concatenation, forward or backward
        total = None
        for f, t, w in zip(self._features(self.normalize(x)), self._targets, self.weights):
            if w == 0.0:
                continue
            term = F.mse_loss(f, t, reduction="sum") * (w * w)
            total = term if total is None else total + term
        return total / self.im_feature.numel()


class MaterialGANOptim:
    """src/materialgan.py:21-153 (checkpoints replaced by seeded weights)."""

    def __init__(self, device, renderer_obj, ckp=None, seed=7):
        self.device = torch.device(device)
        self.renderer_obj = renderer_obj
        self.net_obj = StyleGAN2Generator(renderer_obj.res, 9, seed=seed, device=self.device)
        self.loss_large_feature = VGGLoss(self.device, np.array([1, 1, 4, 8]) / 14)
        self.loss_small_feature = VGGLoss(self.device, np.array([8, 8, 2, 1]) / 19)
        self.noises = None
        self.latent = None
        self._graphs = {}

    def init_from(self, ckp=(), seed=0):
        """Latent W+ from a random z through mapping + truncation (or a saved
        tensor), and the noise maps (src/materialgan.py:58-75)."""
        gen = torch.Generator().manual_seed(seed)
        if len(ckp) == 0:
            z = torch.randn((1, 512), generator=gen).to(self.device)
            with torch.no_grad():
                latent = self.net_obj.truncation(self.net_obj.mapping(z))
        else:
            latent = torch.load(ckp[0], map_location=self.device, weights_only=True)
        self.latent = latent.detach().clone().requires_grad_(True)
        if len(ckp) == 2:
            self.noises = [n.to(self.device).requires_grad_(True)
                           for n in torch.load(ckp[1], map_location=self.device, weights_only=True)]
        else:
            self.noises = [n.requires_grad_(True) for n in self.net_obj.make_noises(gen)]
        self._graphs = {}   # recorded for the previous latent / noise tensors

    def load_targets(self, targets):
        self._graphs = {}
        self.targets = targets
        self.loss_large_feature.load(targets)
        self.loss_small_feature.load(targets)

    def compute_image_loss(self, predicts):
        return F.mse_loss(predicts, self.targets)

    def compute_feature_loss(self, predicts, flag):
        if flag == "L":
            return self.loss_large_feature(predicts)
        if flag == "N":
            return self.loss_small_feature(predicts)
        raise ValueError("a feature-loss flag is needed, 'L' or 'N'")

    def latent_to_textures(self, latent, noises=None):
        t = self.net_obj.synthesis(latent, self.noises if noises is None else noises)
        # channels 0-4 and 6-8 clamped to [-1, 1], roughness (5) to [-0.5, 0.5]:
        # one clamp with per-channel bounds
        if getattr(self, "_clamp_bounds", None) is None or self._clamp_bounds[0].device != t.device:
            lo = torch.full((1, 9, 1, 1), -1.0, device=t.device)
            lo[:, 5] = -0.5
            self._clamp_bounds = (lo, -lo)
        return torch.clamp(t, min=self._clamp_bounds[0], max=self._clamp_bounds[1])

    save_loss = staticmethod(SvbrdfOptim.save_loss)

    # ------------------------------------------------------------ one epoch
    @staticmethod
    def _adam_first_step(params, exp_avg, exp_avg_sq, lr):
        """The step a fresh torch.optim.Adam(params, lr, betas=(0.9, 0.999))
        takes: its foreach implementation on the GPU (torch/optim/adam.py,
        _multi_tensor_adam, step 1, capturable=False) op for op, with the same
        Python scalars, so the same bits -- written out so that the epoch can
        be recorded in a HIP graph."""
        grads = [p.grad for p in params]
        torch._foreach_zero_(exp_avg)
        torch._foreach_zero_(exp_avg_sq)
        torch._foreach_lerp_(exp_avg, grads, 1 - 0.9)
        torch._foreach_mul_(exp_avg_sq, 0.999)
        torch._foreach_addcmul_(exp_avg_sq, grads, grads, 1 - 0.999)
        bc1 = 1 - 0.9 ** 1.0
        bc2 = 1 - 0.999 ** 1.0
        denom = torch._foreach_sqrt(exp_avg_sq)
        torch._foreach_div_(denom, [bc2 ** 0.5] * len(params))
        torch._foreach_add_(denom, 1e-8)
        torch._foreach_addcdiv_(params, exp_avg, denom, [(lr / bc1) * -1] * len(params))

    def _epoch(self, which, lr):
        """One epoch of optim() (a fresh Adam on the latent ("L") or the noise
        maps ("N")); returns the image and feature losses as device scalars."""
        params = [self.latent] if which == "L" else list(self.noises)
        for q in params:
            q.grad = None                                            # optimizer.zero_grad()
        # the epoch's Adam steps only its own parameters: the other kind enters
        # detached, so no gradient is computed for it (the reference's
        # autograd computes one and discards it)
        latent = self.latent if which == "L" else self.latent.detach()
        noises = list(self.noises) if which == "N" else [n.detach() for n in self.noises]
        textures = self.latent_to_textures(latent, noises)
        rendereds = self.renderer_obj.eval(textures)
        loss_image = self.compute_image_loss(rendereds)
        loss_feature = self.compute_feature_loss(rendereds, which) * 0.1
        (loss_image + loss_feature).backward()
        key = ("moments", which)
        if key not in self._graphs:
            self._graphs[key] = ([torch.zeros_like(q) for q in params], [torch.zeros_like(q) for q in params])
        with torch.no_grad():
            self._adam_first_step(params, *self._graphs[key], lr)
        return loss_image.detach(), loss_feature.detach()

    def _graph_epoch(self, which, lr):
        """_epoch replayed from a HIP graph: the first epoch of each kind (and
        learning rate) runs on a side stream as the warm-up the capture needs,
        then the epoch is recorded; later epochs of that kind are one graph
        launch.  The iteration launches ~1,500 kernels, most of them small,
        so this removes the host's launch time from every epoch."""
        key = (which, float(lr))
        if key in self._graphs:
            g, li, lf = self._graphs[key]
            g.replay()
            return li.clone(), lf.clone()
        dev = self.device
        side = torch.cuda.Stream(dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            li, lf = self._epoch(which, lr)
        torch.cuda.current_stream(dev).wait_stream(side)
        # the recorded backward allocates every gradient from the graph's pool
        # and rewrites it on each replay
        for q in [self.latent] + list(self.noises):
            q.grad = None
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            gli, glf = self._epoch(which, lr)
        self._graphs[key] = (g, gli, glf)
        return li, lf

    def optim(self, epochs, lr, svbrdf_obj=None, optim_light=False, save_every=200, graph=True):
        """epochs = (total, l_epochs, n_epochs): l_epochs optimising the latent
        (feature loss "L") then n_epochs the noise ("N"), repeated; a fresh
        Adam each epoch as the reference builds it (src/materialgan.py:
        114-126).  With an ``svbrdf_obj`` that has ``optimize_dir``, the loss
        plot, latent, noise, maps and re-renders are saved after epoch 1,
        every 200th and the last under optimize_dir/tmp/<timestamp>/<epoch>/
        (:103-105, 157-169).  Afterwards ``textures`` holds the maps of the
        final latent and noise, and ``loss`` / ``loss_image`` the last epoch's
        losses (:171-173).  Returns the per-epoch (image, feature) losses.
        graph=True (without optim_light) replays each epoch from a HIP graph
        per kind: the same kernels and Adam step, bit-identical results."""
        total, l_epochs, n_epochs = epochs
        cycle = l_epochs + n_epochs
        tmp_dir = None
        if svbrdf_obj is not None and hasattr(svbrdf_obj, "optimize_dir"):
            stamp = str(datetime.datetime.now()).replace(" ", "-").replace(":", "-").replace(".", "-")
            tmp_dir = pathlib.Path(svbrdf_obj.optimize_dir) / "tmp" / stamp
            tmp_dir.mkdir(parents=True, exist_ok=True)
        light = None
        if optim_light:
            light = (svbrdf_obj.cl[2] if svbrdf_obj is not None else self.renderer_obj.light_pow)
            light = light.detach().to(self.device, torch.float32).reshape(3).clone().requires_grad_(True)
            if svbrdf_obj is not None:
                svbrdf_obj.cl[2] = light
        image_losses, feature_losses = [], []
        use_graph = graph and light is None
        for epoch in range(total):
            which = "L" if (epoch % cycle) < l_epochs else "N"
            if use_graph:
                loss_image, loss_feature = self._graph_epoch(which, lr)
                loss = loss_image + loss_feature
            else:
                params = [self.latent] if which == "L" else list(self.noises)
                if light is not None:
                    params = params + [light]
                    self.renderer_obj.update_light(light)
                optimizer = torch.optim.Adam(params, lr=lr, betas=(0.9, 0.999))
                latent = self.latent if which == "L" else self.latent.detach()
                noises = list(self.noises) if which == "N" else [n.detach() for n in self.noises]
                textures = self.latent_to_textures(latent, noises)
                rendereds = self.renderer_obj.eval(textures)
                loss_image = self.compute_image_loss(rendereds)
                loss_feature = self.compute_feature_loss(rendereds, which) * 0.1
                loss = loss_image + loss_feature
                optimizer.zero_grad()
                loss.backward()
                optimizer.step()
            image_losses.append(loss_image.detach())
            feature_losses.append(loss_feature.detach())
            if tmp_dir is not None and ((epoch + 1) % save_every == 0 or epoch == 0 or epoch == total - 1):
                this_dir = tmp_dir / f"{epoch + 1}"
                this_dir.mkdir(parents=True, exist_ok=True)
                self.save_loss([torch.stack(image_losses).cpu().numpy(), torch.stack(feature_losses).cpu().numpy()],
                               ["image loss", "feature loss"], tmp_dir / "loss.jpg", total)
                torch.save(self.latent.detach(), this_dir / "optim_latent.pt")
                torch.save([n.detach() for n in self.noises], this_dir / "optim_noise.pt")
                with torch.no_grad():
                    saved = self.latent_to_textures(self.latent)
                    if hasattr(svbrdf_obj, "save_textures_th"):
                        svbrdf_obj.save_textures_th(saved, this_dir)
                    if hasattr(svbrdf_obj, "save_images_th"):
                        svbrdf_obj.save_images_th(self.renderer_obj.eval(saved), this_dir)
        if total > 0:
            # the reference leaves the maps of the final latent and noise here
            # (its last epoch is always a save epoch, which re-synthesises them)
            with torch.no_grad():
                self.textures = self.latent_to_textures(self.latent).detach()
            self.loss = float(loss.detach())
            self.loss_image = float(loss_image.detach())
        self.loss_image_list = [float(x) for x in torch.stack(image_losses).cpu()] if image_losses else []
        self.loss_feature_list = [float(x) for x in torch.stack(feature_losses).cpu()] if feature_losses else []
        return self.loss_image_list, self.loss_feature_list
