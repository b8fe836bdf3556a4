"""Convolution layers of the MaterialGAN caller's networks on the matrix-core
kernels of libsvbrdf_hip.so (csrc/conv.hip).

* :func:`conv3x3` -- a 3x3 convolution (stride 1, zero padding 1) with the
  fused epilogue of the reference's blocks: ``act(conv + noise_strength *
  noise + bias)``, act = identity, ReLU (VGG, src/descriptor.py:7-79) or
  leaky ReLU x sqrt(2) (ModulateConvBlock,
  src/higan_models/stylegan2_generator_network.py:738-755).  Differentiable
  with respect to the input, the weights and the noise map: the forward and
  dL/dx are the HIP kernel (dL/dx is the convolution of dL/dy with the
  flipped, transposed weights; where only dL/dx is wanted the kernel applies
  the activation's derivative to dL/dy itself, :func:`conv3x3_input_grad`),
  and so is dL/dweights (:func:`conv3x3_wgrad`, an MFMA GEMM over the pixels
  with the nine taps sharing the dL/dy operand).
* :func:`upsample2` -- x2 upsampling by zero insertion and the [1, 3, 3, 1]
  filter (UpsamplingLayer, stylegan2_generator_network.py:419-479); forward
  and adjoint are HIP kernels.
* :class:`Conv3x3` -- a frozen 3x3 layer (VGG): weights packed once, in both
  layouts.
"""
import torch

from . import _lib

ACT = {"none": _lib.ACT_NONE, "relu": _lib.ACT_RELU, "lrelu": _lib.ACT_LRELU}


def _check(t, name):
    if t.device.type != "cuda" or t.dtype != torch.float32:
        raise ValueError(f"{name} must be a float32 CUDA tensor (there is no CPU path), got {t.dtype} on {t.device}")


def pack3x3(weight, style=None, scale=1.0, demodulate=False, transpose=False):
    """[Cout, Cin, 3, 3] -> the kernel's layout ([9, Cin, Cout], or the input-
    gradient layout [9, Cout, Cin] with flipped taps when transpose), times
    scale * style[ci] and demodulated per output channel when asked."""
    w = weight.detach().contiguous()
    _check(w, "weight")
    cout, cin = int(w.shape[0]), int(w.shape[1])
    if tuple(w.shape[2:]) != (3, 3):
        raise ValueError(f"weight must be [Cout, Cin, 3, 3], got {tuple(w.shape)}")
    st = None
    if style is not None:
        st = style.detach().reshape(-1).contiguous()
        _check(st, "style")
        if st.numel() != cin:
            raise ValueError(f"style must have {cin} values, got {st.numel()}")
    out = torch.empty((9, cout, cin) if transpose else (9, cin, cout), dtype=torch.float32, device=w.device)
    _lib.call("svbrdf_conv3x3_pack", _lib.ptr(w), cout, cin, _lib.ptr(st), float(scale), int(bool(demodulate)),
              int(bool(transpose)), _lib.ptr(out), _lib.stream_handle(w.device))
    return out


def conv3x3_packed(x, packed, cout, bias=None, noise=None, noise_strength=0.0, act="none"):
    """The kernel itself on packed weights: x [N, Cin, H, W] -> [N, Cout, H, W]."""
    x = x.contiguous()
    _check(x, "x")
    n, cin, h, w = (int(s) for s in x.shape)
    if tuple(packed.shape) != (9, cin, cout):
        raise ValueError(f"packed weights {tuple(packed.shape)} do not match Cin={cin}, Cout={cout}")
    if bias is not None:
        bias = bias.detach().reshape(-1).contiguous()
        _check(bias, "bias")
    if noise is not None:
        noise = noise.detach().reshape(-1).contiguous()
        _check(noise, "noise")
        if noise.numel() != h * w:
            raise ValueError(f"noise must have {h * w} values, got {noise.numel()}")
    y = torch.empty((n, cout, h, w), dtype=torch.float32, device=x.device)
    # K-split partial sums of small images (torch's caching allocator makes
    # this a free-list lookup after the first call)
    ws_bytes = int(_lib.query("svbrdf_conv3x3_workspace_bytes", n, cin, h, w, cout))
    ws = torch.empty(ws_bytes // 4, dtype=torch.float32, device=x.device) if ws_bytes else None
    _lib.call("svbrdf_conv3x3", _lib.ptr(x), n, cin, h, w, _lib.ptr(packed), cout, _lib.ptr(bias), _lib.ptr(noise),
              float(noise_strength), ACT[act], _lib.ptr(y), _lib.ptr(ws), ws_bytes, _lib.stream_handle(x.device))
    return y


def conv3x3_input_grad(grad_y, y, packed_t, cin, act):
    """dL/dx of y = act(conv3x3(x, w) + ...) from dL/dy and y, with the
    flipped, transposed weights of pack3x3(w, transpose=True): the kernel
    gates dL/dy by the activation's derivative while it stages it."""
    g = grad_y.contiguous()
    _check(g, "grad_y")
    y = y.contiguous()
    _check(y, "y")
    n, cout, h, w = (int(s) for s in g.shape)
    if tuple(y.shape) != tuple(g.shape):
        raise ValueError(f"y {tuple(y.shape)} and grad_y {tuple(g.shape)} differ")
    if tuple(packed_t.shape) != (9, cout, cin):
        raise ValueError(f"packed weights {tuple(packed_t.shape)} do not match Cout={cout}, Cin={cin}")
    gx = torch.empty((n, cin, h, w), dtype=torch.float32, device=g.device)
    ws_bytes = int(_lib.query("svbrdf_conv3x3_workspace_bytes", n, cout, h, w, cin))
    ws = torch.empty(ws_bytes // 4, dtype=torch.float32, device=g.device) if ws_bytes else None
    _lib.call("svbrdf_conv3x3_input_grad", _lib.ptr(g), _lib.ptr(y), ACT[act], n, cout, h, w, _lib.ptr(packed_t), cin,
              _lib.ptr(gx), _lib.ptr(ws), ws_bytes, _lib.stream_handle(g.device))
    return gx


def conv3x3_wgrad(x, grad_z, cout):
    """dL/dw [cout, Cin, 3, 3] of a 3x3 convolution (padding 1) from its input
    x [N, Cin, H, W] and grad_z [N, cout, H, W], the gradient with respect to
    its output before bias and activation (torch.nn.grad.conv2d_weight)."""
    x = x.contiguous()
    g = grad_z.contiguous()
    _check(x, "x")
    _check(g, "grad_z")
    n, cin, h, w = (int(s) for s in x.shape)
    if tuple(g.shape) != (n, cout, h, w):
        raise ValueError(f"grad_z {tuple(g.shape)} does not match x {tuple(x.shape)} and Cout={cout}")
    gw = torch.empty((cout, cin, 3, 3), dtype=torch.float32, device=x.device)
    ws_bytes = int(_lib.query("svbrdf_conv3x3_wgrad_workspace_bytes", n, cin, h, w, cout))
    ws = torch.empty(ws_bytes // 4, dtype=torch.float32, device=x.device) if ws_bytes else None
    _lib.call("svbrdf_conv3x3_wgrad", _lib.ptr(x), _lib.ptr(g), n, cin, h, w, cout, _lib.ptr(gw), _lib.ptr(ws),
              ws_bytes, _lib.stream_handle(x.device))
    return gw


_ACT_SLOPES = {"relu": (1.0, 0.0), "lrelu": (1.4142135623730951, 0.2 * 1.4142135623730951)}


def _act_grad(grad_y, y, act):
    """dL/d(pre-activation) from dL/dy and y (the activations keep the sign):
    one HIP kernel, grad_y * (y > 0 ? pos : neg)."""
    if act not in _ACT_SLOPES:
        return grad_y
    g, yy = grad_y.contiguous(), y.contiguous()
    out = torch.empty_like(g)
    pos, neg = _ACT_SLOPES[act]
    _lib.call("svbrdf_act_grad", _lib.ptr(g), _lib.ptr(yy), g.numel(), pos, neg, _lib.ptr(out),
              _lib.stream_handle(g.device))
    return out


class Conv3x3Function(torch.autograd.Function):
    """y = act(conv3x3(x, weight) + noise_strength * noise + bias)."""

    @staticmethod
    def forward(ctx, x, weight, bias, noise, noise_strength, act):
        packed = pack3x3(weight)
        y = conv3x3_packed(x, packed, int(weight.shape[0]), bias, noise, noise_strength, act)
        ctx.save_for_backward(x, weight, noise, y)
        ctx.noise_strength, ctx.act = float(noise_strength), act
        return y

    @staticmethod
    def backward(ctx, grad_y):
        x, weight, noise, y = ctx.saved_tensors
        gx = gw = gnoise = None
        need_gz = ctx.needs_input_grad[1] or (ctx.needs_input_grad[3] and noise is not None)
        if not need_gz:
            # only dL/dx: the gate is applied inside the convolution
            if ctx.needs_input_grad[0]:
                gx = conv3x3_input_grad(grad_y, y, pack3x3(weight, transpose=True), int(weight.shape[1]), ctx.act)
            return gx, None, None, None, None, None
        gz = _act_grad(grad_y.contiguous(), y, ctx.act).contiguous()
        if ctx.needs_input_grad[0]:
            gx = conv3x3_packed(gz, pack3x3(weight, transpose=True), int(weight.shape[1]))
        if ctx.needs_input_grad[1]:
            gw = conv3x3_wgrad(x, gz, int(weight.shape[0]))
        if ctx.needs_input_grad[3] and noise is not None:
            gnoise = (ctx.noise_strength * gz.sum(dim=(0, 1))).reshape(noise.shape)
        return gx, gw, None, gnoise, None, None


def conv3x3(x, weight, bias=None, noise=None, noise_strength=0.0, act="none"):
    if act not in ACT:
        raise ValueError(f"act must be one of {sorted(ACT)}")
    return Conv3x3Function.apply(x, weight, bias, noise, noise_strength, act)


def modulation_grad(weight, style, scale, demodulate, grad_weight):
    """dL/dstyle [Cin] of the weights pack3x3(weight, style, scale,
    demodulate) packs, from dL/d(those weights) in [Cout, Cin, 3, 3]."""
    w = weight.detach().contiguous()
    g = grad_weight.contiguous()
    st = style.detach().reshape(-1).contiguous()
    for t, name in ((w, "weight"), (g, "grad_weight"), (st, "style")):
        _check(t, name)
    cout, cin = int(w.shape[0]), int(w.shape[1])
    if tuple(g.shape) != tuple(w.shape) or st.numel() != cin:
        raise ValueError(f"grad_weight {tuple(g.shape)} / style {st.numel()} do not match weight {tuple(w.shape)}")
    ws = torch.empty(2 * cout, dtype=torch.float32, device=w.device)
    out = torch.empty(cin, dtype=torch.float32, device=w.device)
    _lib.call("svbrdf_conv3x3_modulation_grad", _lib.ptr(w), cout, cin, _lib.ptr(st), float(scale),
              int(bool(demodulate)), _lib.ptr(g), _lib.ptr(out), _lib.ptr(ws), ws.numel() * 4,
              _lib.stream_handle(w.device))
    return out


class ModConv3x3Function(torch.autograd.Function):
    """The modulated 3x3 convolution of a StyleGAN2 block (batch 1):
    y = act(conv3x3(x, w') + noise_strength * noise + bias) with
    w'[co][ci] = scale style[ci] w[co][ci], divided per output channel by
    sqrt(sum w'^2 + 1e-8) when demodulate (ModulateConvBlock.forward,
    src/higan_models/stylegan2_generator_network.py:694-755).  The modulation
    and demodulation happen inside the weight packing; the backward gives
    dL/dx, dL/dstyle (through svbrdf_conv3x3_modulation_grad) and dL/dnoise.
    The base weight w is frozen (the caller optimises latents and noise)."""

    @staticmethod
    def forward(ctx, x, style, weight, scale, demodulate, bias, noise, noise_strength, act):
        if weight.requires_grad:
            raise ValueError("ModConv3x3Function keeps the base weight frozen; use conv3x3 with explicitly "
                             "modulated weights to train it")
        packed = pack3x3(weight, style, scale, demodulate)
        y = conv3x3_packed(x, packed, int(weight.shape[0]), bias, noise, noise_strength, act)
        ctx.save_for_backward(x, style, weight, noise, y)
        ctx.scale, ctx.demodulate = float(scale), bool(demodulate)
        ctx.noise_strength, ctx.act = float(noise_strength), act
        return y

    @staticmethod
    def backward(ctx, grad_y):
        x, style, weight, noise, y = ctx.saved_tensors
        gx = gstyle = gnoise = None
        cin = int(weight.shape[1])
        need_gx, need_gs = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        need_gn = ctx.needs_input_grad[6] and noise is not None
        if need_gx:
            packed_t = pack3x3(weight, style, ctx.scale, ctx.demodulate, transpose=True)
        if not (need_gs or need_gn):
            if need_gx:
                gx = conv3x3_input_grad(grad_y, y, packed_t, cin, ctx.act)
            return gx, None, None, None, None, None, None, None, None
        gz = _act_grad(grad_y.contiguous(), y, ctx.act).contiguous()
        if need_gx:
            gx = conv3x3_packed(gz, packed_t, cin)
        if need_gs:
            gw = conv3x3_wgrad(x, gz, int(weight.shape[0]))
            gstyle = modulation_grad(weight, style, ctx.scale, ctx.demodulate, gw).reshape(style.shape)
        if need_gn:
            gnoise = (ctx.noise_strength * gz.sum(dim=(0, 1))).reshape(noise.shape)
        return gx, gstyle, None, None, None, None, gnoise, None, None


def modulated_conv3x3(x, style, weight, scale, demodulate=True, bias=None, noise=None, noise_strength=0.0,
                      act="none"):
    if act not in ACT:
        raise ValueError(f"act must be one of {sorted(ACT)}")
    if x.shape[0] != 1:
        raise ValueError("modulated_conv3x3 takes one sample (the style is per input channel)")
    return ModConv3x3Function.apply(x, style, weight, scale, demodulate, bias, noise, noise_strength, act)


class Upsample2Function(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        x = x.contiguous()
        _check(x, "x")
        n, c, h, w = (int(s) for s in x.shape)
        y = torch.empty((n, c, 2 * h, 2 * w), dtype=torch.float32, device=x.device)
        _lib.call("svbrdf_upsample2_blur", _lib.ptr(x), n * c, h, w, _lib.ptr(y), _lib.stream_handle(x.device))
        return y

    @staticmethod
    def backward(ctx, grad_y):
        g = grad_y.contiguous()
        n, c, h2, w2 = (int(s) for s in g.shape)
        gx = torch.empty((n, c, h2 // 2, w2 // 2), dtype=torch.float32, device=g.device)
        _lib.call("svbrdf_upsample2_blur_adjoint", _lib.ptr(g), n * c, h2 // 2, w2 // 2, _lib.ptr(gx),
                  _lib.stream_handle(g.device))
        return gx


def upsample2(x):
    return Upsample2Function.apply(x)


class AvgPool2Function(torch.autograd.Function):
    """F.avg_pool2d(x, 2) (src/descriptor.py:15-18 puts AvgPool2d(2) in
    place of VGG's max pooling) on the HIP kernels: same bits forward, and
    the adjoint writes dL/dy / 4 to each window."""

    @staticmethod
    def forward(ctx, x):
        x = x.contiguous()
        _check(x, "x")
        n, c, h, w = (int(s) for s in x.shape)
        y = torch.empty((n, c, h // 2, w // 2), dtype=torch.float32, device=x.device)
        _lib.call("svbrdf_avgpool2", _lib.ptr(x), n * c, h, w, _lib.ptr(y), _lib.stream_handle(x.device))
        ctx.shape = (n, c, h, w)
        return y

    @staticmethod
    def backward(ctx, grad_y):
        g = grad_y.contiguous()
        n, c, h, w = ctx.shape
        gx = torch.empty((n, c, h, w), dtype=torch.float32, device=g.device)
        _lib.call("svbrdf_avgpool2_adjoint", _lib.ptr(g), n * c, h, w, _lib.ptr(gx), _lib.stream_handle(g.device))
        return gx


def avgpool2(x):
    return AvgPool2Function.apply(x)


class _FrozenConvFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, layer):
        y = conv3x3_packed(x, layer.packed, layer.cout, layer.bias, None, 0.0, layer.act)
        ctx.save_for_backward(y)
        ctx.layer = layer
        return y

    @staticmethod
    def backward(ctx, grad_y):
        (y,) = ctx.saved_tensors
        layer = ctx.layer
        return conv3x3_input_grad(grad_y, y, layer.packed_t, layer.cin, layer.act), None


class Conv3x3(torch.nn.Module):
    """A 3x3 layer with frozen weights [Cout, Cin, 3, 3] and bias [Cout],
    followed by `act`: packed once in the forward and input-gradient layouts."""

    def __init__(self, weight, bias=None, act="relu"):
        super().__init__()
        if act not in ACT:
            raise ValueError(f"act must be one of {sorted(ACT)}")
        self.cout, self.cin = int(weight.shape[0]), int(weight.shape[1])
        self.weight = weight.detach().contiguous()
        self.bias = None if bias is None else bias.detach().contiguous()
        self.act = act
        self.packed = pack3x3(self.weight)
        self.packed_t = pack3x3(self.weight, transpose=True)

    def forward(self, x):
        return _FrozenConvFunction.apply(x, self)
