"""The matrix-core 3x3 convolution (csrc/conv.hip through svbrdf_amd.convnet)
against torch's F.conv2d on the same weights in FP32 and FP64, and the
networks built on it (the MaterialGAN generator's synthesis pass, the two VGG
stacks) against their torch restatements (tests/convnet_ref.py).

Forward criterion: every output within 2e-5 of the FP64 result relative to
the convolution of |x| with |w| (the scale of a dot product's rounding
error), plus 1e-7.  Gradients: norm-wise within 1e-5 of FP64.  Through a
deep stack of ReLU / leaky-ReLU layers the FP32 forward lands on the other
side of a kink from FP64 at a handful of pre-activations within rounding of
0 (torch's FP32 convolutions as well), and each such branch flip changes the
gradient by a finite step; there the FP64 gradient is taken with the FP32
forward's branches, and the flips themselves are bounded (at most 1e-6 of
the activations, each within 1e-4 of its layer's largest pre-activation)."""
import math
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from tests import convnet_ref as R

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")

SHAPES = [  # n, cin, cout, h, w, act, noise, bias
    (1, 512, 512, 4, 4, "lrelu", True, True),
    (1, 512, 512, 16, 16, "lrelu", True, True),
    (1, 256, 128, 64, 64, "lrelu", True, True),
    (9, 3, 64, 32, 32, "relu", False, True),
    (2, 64, 64, 37, 29, "none", False, False),
    (1, 13, 70, 5, 7, "none", False, True),
    (3, 128, 256, 24, 40, "relu", False, True),
    # at most 8 output channels: the direct kernel
    (2, 64, 3, 33, 70, "none", False, False),
    (1, 20, 7, 9, 66, "lrelu", True, True),
    (4, 8, 8, 17, 5, "relu", False, True),
]


def _inputs(n, cin, cout, h, w, seed):
    g = torch.Generator().manual_seed(seed)
    x = torch.randn((n, cin, h, w), generator=g)
    wt = torch.randn((cout, cin, 3, 3), generator=g) / np.sqrt(cin * 9)
    b = torch.randn(cout, generator=g) * 0.1
    nz = torch.randn((1, 1, h, w), generator=g)
    return x.to(DEV), wt.to(DEV), b.to(DEV), nz.to(DEV)


def _check_forward(y, x, wt, b, nz, ns, act, label):
    with torch.no_grad():
        y64 = R.conv3x3_ref(x.double(), wt.double(), b, nz, ns, act)
        y32 = R.conv3x3_ref(x, wt, b, nz, ns, act)
        scale = F.conv2d(x.double().abs(), wt.double().abs(), padding=1) + (abs(ns) * nz.double().abs() if nz is not None else 0)
    err = (y.double() - y64).abs()
    bound = 2e-5 * scale + 1e-7
    ratio = float((err / bound).max())
    assert ratio <= 1.0, f"{label}: error {float(err.max()):.3e}, worst at {ratio:.2f} of the bound"
    return float((err / (scale + 1e-30)).max()), float(((y32.double() - y64).abs() / (scale + 1e-30)).max())


@pytest.mark.parametrize("shape", SHAPES, ids=[f"{s[0]}x{s[1]}->{s[2]}@{s[3]}x{s[4]}" for s in SHAPES])
def test_conv3x3_forward(pkg, shape):
    from svbrdf_amd import convnet
    n, cin, cout, h, w, act, use_noise, use_bias = shape
    x, wt, b, nz = _inputs(n, cin, cout, h, w, seed=cin + cout + h)
    b = b if use_bias else None
    nz = nz if use_noise else None
    ns = 0.37 if use_noise else 0.0
    y = convnet.conv3x3(x, wt, b, nz, ns, act)
    # the K slices of small images are added in a fixed order: repeatable bits
    assert torch.equal(y, convnet.conv3x3(x, wt, b, nz, ns, act))
    torch.cuda.synchronize()
    assert y.shape == (n, cout, h, w)
    e, e32 = _check_forward(y, x, wt, b, nz, ns, act, str(shape))
    print(f"conv {shape}: max error / scale {e:.2e} (torch FP32: {e32:.2e})")


def test_conv3x3_gradients(pkg):
    """dL/dx (the same kernel on the flipped, transposed weights), dL/dw and
    dL/dnoise through the leaky-ReLU epilogue against FP64 autograd."""
    from svbrdf_amd import convnet
    x, wt, b, nz = _inputs(2, 96, 160, 20, 28, seed=5)
    gy = torch.randn((2, 160, 20, 28), generator=torch.Generator().manual_seed(6)).to(DEV)
    xs, ws, ns_ = x.clone().requires_grad_(True), wt.clone().requires_grad_(True), nz.clone().requires_grad_(True)
    (convnet.conv3x3(xs, ws, b, ns_, 0.3, "lrelu") * gy).sum().backward()
    x64, w64, n64 = (t.double().requires_grad_(True) for t in (x, wt, nz))
    (R.conv3x3_ref(x64, w64, b, n64, 0.3, "lrelu") * gy.double()).sum().backward()
    for name, g, g64 in (("x", xs.grad, x64.grad), ("w", ws.grad, w64.grad), ("noise", ns_.grad, n64.grad)):
        rel = float((g.double() - g64).norm() / g64.norm())
        print(f"conv gradient d{name}: {rel:.2e}")
        assert rel <= 1e-5, (name, rel)


GATED = [(2, 64, 96, 20, 28), (1, 512, 512, 8, 8), (3, 3, 48, 33, 40)]   # n, cin, cout, h, w of the layer


@pytest.mark.parametrize("act", ["relu", "lrelu", "none"])
@pytest.mark.parametrize("shape", GATED, ids=["mfma", "k-split", "direct"])
def test_conv3x3_input_grad_gated(pkg, act, shape):
    """svbrdf_conv3x3_input_grad (dL/dy gated by the activation's derivative
    while the kernel stages it) against FP64 autograd through the same
    activation branches as the FP32 forward, and against the two-pass path
    (torch gate, then the plain input-gradient convolution): the MFMA
    kernel, its K split and the direct kernel (a 3-channel input)."""
    from svbrdf_amd import convnet
    n, cin, cout, h, w = shape
    x, wt, b, _ = _inputs(n, cin, cout, h, w, seed=cin + 3 * cout)
    y = convnet.conv3x3(x, wt, b, None, 0.0, act)
    gy = torch.randn(y.shape, generator=torch.Generator().manual_seed(12)).to(DEV)
    packed_t = convnet.pack3x3(wt, transpose=True)
    gx = convnet.conv3x3_input_grad(gy, y, packed_t, cin, act)
    two_pass = convnet.conv3x3_packed(convnet._act_grad(gy, y, act).contiguous(), packed_t, cin)
    x64 = x.double().requires_grad_(True)
    out = R.conv3x3_ref(x64, wt.double(), b, None, 0.0, act, mask=(y > 0) if act != "none" else None)
    (out * gy.double()).sum().backward()
    torch.cuda.synchronize()
    rel = float((gx.double() - x64.grad).norm() / x64.grad.norm())
    rel2 = float((gx - two_pass).norm() / two_pass.norm())
    print(f"gated dL/dx {shape} {act}: {rel:.2e} vs FP64, {rel2:.2e} vs the two-pass path")
    assert rel <= 1e-5, rel
    assert rel2 <= 1e-6, rel2


WGRAD = [  # n, cin, cout, h, w
    (2, 96, 160, 20, 28),
    (1, 512, 512, 4, 4),
    (1, 128, 128, 256, 256),
    (3, 13, 70, 5, 7),
    (1, 3, 64, 33, 17),
    (4, 64, 3, 16, 16),
    (1, 256, 512, 64, 64),
]


@pytest.mark.parametrize("shape", WGRAD, ids=[f"{s[0]}x{s[1]}->{s[2]}@{s[3]}x{s[4]}" for s in WGRAD])
def test_conv3x3_weight_gradient(pkg, shape):
    """svbrdf_conv3x3_wgrad against torch.nn.grad.conv2d_weight in FP64: every
    element within 2e-5 of the FP64 value relative to the same sum over |x|
    and |dL/dz| (the scale of its rounding error), plus 1e-7; the pixel-sum
    slices of the large images are added in a fixed order (repeatable bits)."""
    from svbrdf_amd import convnet
    n, cin, cout, h, w = shape
    g = torch.Generator().manual_seed(cin * 7 + cout)
    x = torch.randn((n, cin, h, w), generator=g).to(DEV)
    gz = torch.randn((n, cout, h, w), generator=g).to(DEV)
    gw = convnet.conv3x3_wgrad(x, gz, cout)
    assert torch.equal(gw, convnet.conv3x3_wgrad(x, gz, cout))
    with torch.no_grad():
        ref = torch.nn.grad.conv2d_weight(x.double(), (cout, cin, 3, 3), gz.double(), padding=1)
        scale = torch.nn.grad.conv2d_weight(x.double().abs(), (cout, cin, 3, 3), gz.double().abs(), padding=1)
        ref32 = torch.nn.grad.conv2d_weight(x, (cout, cin, 3, 3), gz, padding=1)
    torch.cuda.synchronize()
    err = (gw.double() - ref).abs()
    ratio = float((err / (2e-5 * scale + 1e-7)).max())
    e32 = float(((ref32.double() - ref).abs() / (scale + 1e-30)).max())
    print(f"wgrad {shape}: max error / scale {float((err / (scale + 1e-30)).max()):.2e} (MIOpen FP32 {e32:.2e})")
    assert ratio <= 1.0, (shape, ratio)


def test_conv3x3_refuses_a_short_workspace(pkg):
    """A shape that splits the channel sum needs the queried workspace; one
    float short is refused before any launch."""
    from svbrdf_amd import _lib
    from svbrdf_amd import convnet
    x, wt, _, _ = _inputs(1, 512, 512, 4, 4, seed=1)
    packed = convnet.pack3x3(wt)
    y = torch.empty((1, 512, 4, 4), device=DEV)
    need = int(_lib.query("svbrdf_conv3x3_workspace_bytes", 1, 512, 4, 4, 512))
    assert need > 0
    ws = torch.empty(need // 4, device=DEV)
    with pytest.raises(ValueError, match="workspace"):
        _lib.call("svbrdf_conv3x3", _lib.ptr(x), 1, 512, 4, 4, _lib.ptr(packed), 512, None, None, 0.0, 0,
                  _lib.ptr(y), _lib.ptr(ws), need - 4, _lib.stream_handle(DEV))


def test_pack_folds_style_and_demodulates(pkg):
    from svbrdf_amd import convnet
    g = torch.Generator().manual_seed(3)
    w = torch.randn((40, 24, 3, 3), generator=g).to(DEV)
    style = (torch.rand(24, generator=g) + 0.5).to(DEV)
    p = convnet.pack3x3(w, style, scale=0.07, demodulate=True)
    ref = w.double() * 0.07 * style.double().view(1, -1, 1, 1)
    ref = ref * torch.rsqrt((ref * ref).sum(dim=(1, 2, 3), keepdim=True) + 1e-8)
    got = p.double().reshape(3, 3, 24, 40).permute(3, 2, 0, 1)
    assert float((got - ref).abs().max()) <= 1e-6 * float(ref.abs().max())
    pt = convnet.pack3x3(w, transpose=True)
    got_t = pt.double().reshape(3, 3, 40, 24).permute(2, 3, 0, 1).flip(2, 3)
    assert torch.equal(got_t, w.double())


def test_upsample2_and_adjoint(pkg):
    from svbrdf_amd import convnet
    g = torch.Generator().manual_seed(8)
    x = torch.randn((2, 5, 9, 13), generator=g).to(DEV)
    gy = torch.randn((2, 5, 18, 26), generator=g).to(DEV)
    xs = x.clone().requires_grad_(True)
    y = convnet.upsample2(xs)
    (y * gy).sum().backward()
    x64 = x.double().requires_grad_(True)
    y64 = R.upsample2_ref(x64)
    (y64 * gy.double()).sum().backward()
    assert float((y.detach().double() - y64.detach()).abs().max()) <= 1e-6
    assert float((xs.grad.double() - x64.grad).abs().max()) <= 1e-5
    # unit DC gain away from the border
    up = convnet.upsample2(torch.ones(1, 2, 8, 8, device=DEV))
    assert torch.allclose(up[:, :, 1:-1, 1:-1], torch.ones_like(up[:, :, 1:-1, 1:-1]))


@pytest.mark.parametrize("shape", [(9, 64, 256, 256), (2, 3, 7, 10), (1, 5, 9, 9), (1, 1, 2, 2)])
def test_avgpool2_matches_torch_bit_for_bit(pkg, shape):
    """svbrdf_avgpool2 and its adjoint against F.avg_pool2d(x, 2) and its
    autograd: the same bits (window summed row by row, then / 4), odd sizes
    dropping their last row / column."""
    from svbrdf_amd import convnet
    x = torch.randn(shape, generator=torch.Generator().manual_seed(sum(shape))).to(DEV)
    xa, xb = x.clone().requires_grad_(True), x.clone().requires_grad_(True)
    ya, yb = convnet.avgpool2(xa), F.avg_pool2d(xb, 2)
    assert torch.equal(ya, yb)
    gy = torch.randn(ya.shape, generator=torch.Generator().manual_seed(3)).to(DEV)
    ya.backward(gy)
    yb.backward(gy)
    assert torch.equal(xa.grad, xb.grad)


def _masks_of(layers):
    """Forward hooks recording each layer's activation branch (output > 0,
    the sign of the pre-activation); returns (masks, remove)."""
    masks = []
    handles = [m.register_forward_hook(lambda mod, inp, out: masks.append(out.detach() > 0)) for m in layers]
    return masks, lambda: [h.remove() for h in handles]


def _check_kinks(masks, pre64, label):
    """The activation branches of the FP32 forward may differ from FP64's only
    at pre-activations within rounding of 0: at most 1e-6 of them (and 16),
    each within 1e-4 of its layer's largest |pre-activation|."""
    flips = total = 0
    for m, pre in zip(masks, pre64):
        d = m != (pre > 0)
        total += d.numel()
        if d.any():
            flips += int(d.sum())
            near = float(pre[d].abs().max() / pre.abs().max())
            assert near <= 1e-4, f"{label}: a branch flip at |pre| = {near:.2e} of the layer's largest"
    assert flips <= max(16, 1e-6 * total), f"{label}: {flips} branch flips in {total} activations"
    return flips, total


def _gradient_check(g, g64m, g64, g32, label):
    """g against the FP64 gradient taken with the same activation branches
    (g64m); the distance to the plain FP64 gradient (g64), all of it from the
    kinks, is reported beside torch FP32's."""
    e = float((g.double() - g64m).norm() / g64m.norm())
    e_all = float((g.double() - g64).norm() / g64.norm())
    e32 = float((g32.double() - g64).norm() / g64.norm())
    print(f"{label}: |g - g64(same branches)|/|g64| {e:.2e}; |g - g64| {e_all:.2e} (torch FP32 {e32:.2e})")
    assert e <= 1e-5, (label, e)


@pytest.mark.parametrize("demod,cin,cout,side", [(True, 64, 32, 16), (False, 32, 64, 8), (True, 512, 512, 4),
                                                (True, 128, 256, 32)])
@pytest.mark.parametrize("act,style_grad", [("none", True), ("lrelu", True), ("lrelu", False)])
def test_modulated_conv3x3_against_torch(pkg, demod, cin, cout, side, act, style_grad):
    """modulated_conv3x3 -- the style modulation and demodulation inside the
    HIP weight packing, dL/dstyle from svbrdf_conv3x3_modulation_grad --
    against the torch restatement of ModulateConvBlock (weights scaled by the
    style, demodulated, then F.conv2d + noise + bias) in FP64: the output,
    dL/dx, dL/dstyle and dL/dnoise, also without a style gradient (the noise
    epochs).  With the leaky ReLU the FP64 side takes its branches from the
    GPU output's signs."""
    from svbrdf_amd import convnet
    g = torch.Generator().manual_seed(cin + cout + side)
    x = torch.randn(1, cin, side, side, generator=g)
    style = torch.randn(1, cin, generator=g) * 0.5 + 1.0
    w = torch.randn(cout, cin, 3, 3, generator=g)
    bias = torch.randn(cout, generator=g) * 0.1
    noise = torch.randn(1, 1, side, side, generator=g)
    probe = torch.randn(1, cout, side, side, generator=g)
    scale = 1.0 / math.sqrt(cin * 9)

    def ref(xx, ss, nn, mask):
        wt = w.double() * scale * ss.view(1, -1, 1, 1)
        if demod:
            wt = wt * torch.rsqrt((wt * wt).sum(dim=(1, 2, 3), keepdim=True) + 1e-8)
        return R.conv3x3_ref(xx, wt, bias.double(), nn, 0.3, act, mask)

    xs, ns = x.to(DEV).requires_grad_(True), noise.to(DEV).requires_grad_(True)
    ss = style.to(DEV).requires_grad_(style_grad)
    y = convnet.modulated_conv3x3(xs, ss, w.to(DEV), scale, demod, bias.to(DEV), ns, 0.3, act)
    (y * probe.to(DEV)).sum().backward()
    torch.cuda.synchronize()
    mask = (y.detach() > 0).cpu() if act == "lrelu" else None
    x64, s64, n64 = (t.double().requires_grad_(True) for t in (x, style, noise))
    y64 = ref(x64, s64, n64, mask)
    (y64 * probe.double()).sum().backward()
    ey = float((y.detach().cpu().double() - y64.detach()).abs().max() / y64.abs().max())
    assert ey <= 1e-5, ey
    checks = [(xs.grad, x64.grad, "dL/dx"), (ns.grad, n64.grad, "dL/dnoise")]
    if style_grad:
        checks.append((ss.grad, s64.grad, "dL/dstyle"))
    else:
        assert ss.grad is None
    for got, want, name in checks:
        e = float((got.cpu().double() - want).norm() / want.norm())
        print(f"{name}: relative error {e:.2e}")
        assert e <= 1e-5, (name, e)


@pytest.mark.parametrize("act", ["relu", "lrelu"])
@pytest.mark.parametrize("n", [4096, 4099])
def test_activation_gradient_kernel_is_torch_bitwise(pkg, act, n):
    """svbrdf_act_grad (four elements per thread, and the one-element tail
    path for odd counts) against torch's dL/dy * act'(y) from y, bit for bit,
    with NaN, infinite and exactly-zero y and dL/dy among the values."""
    from svbrdf_amd import convnet
    g = torch.Generator().manual_seed(n)
    gy = torch.randn(n, generator=g)
    y = torch.randn(n, generator=g)
    y[:3] = torch.tensor([0.0, float("nan"), float("inf")])
    gy[3:6] = torch.tensor([float("inf"), -0.0, float("nan")])
    gy, y = gy.to(DEV), y.to(DEV)
    ref = gy * (y > 0) if act == "relu" else gy * torch.where(y > 0, 1.4142135623730951, 0.2 * 1.4142135623730951)
    out = convnet._act_grad(gy, y, act)
    torch.cuda.synchronize()
    assert torch.equal(torch.isnan(out), torch.isnan(ref))
    ok = ~torch.isnan(ref)
    assert torch.equal(out[ok].view(torch.int32), ref[ok].view(torch.int32))


def test_generator_synthesis_and_gradients(pkg):
    """The 256^2 StyleGAN2-shaped synthesis on the HIP convolutions against its
    torch restatement in FP32 and FP64: the 9-channel output, and the
    gradients with respect to the W+ latent and the 14 noise maps (what
    MaterialGANOptim optimises), the latter against FP64 with the same
    leaky-ReLU branches (tests/test_gpu_convnet.py _check_kinks bounds where
    the branches may differ)."""
    from svbrdf_amd import materialgan as MG
    gen = MG.StyleGAN2Generator(256, 9, seed=1, device=DEV)
    gz = torch.Generator().manual_seed(2)
    wp = gen.truncation(gen.mapping(torch.randn(1, 512, generator=gz).to(DEV)))
    noises = gen.make_noises(torch.Generator().manual_seed(3))
    probe = torch.randn((1, 9, 256, 256), generator=torch.Generator().manual_seed(4)).to(DEV)

    def run(fn, dtype):
        w = wp.detach().to(dtype).requires_grad_(True)
        ns = [n.detach().to(dtype).requires_grad_(True) for n in noises]
        out = fn(w, ns)
        (out * probe.to(dtype)).sum().backward()
        return out.detach(), w.grad, torch.cat([n.grad.flatten() for n in ns])

    masks, remove = _masks_of(gen.convs)
    out, gw, gn = run(gen.synthesis, torch.float32)
    remove()
    pre64 = []
    out64, gw64, gn64 = run(lambda w, n: R.synthesis_ref(gen, w, n, record=pre64), torch.float64)
    _, gw64m, gn64m = run(lambda w, n: R.synthesis_ref(gen, w, n, masks=masks), torch.float64)
    out32, gw32, gn32 = run(lambda w, n: R.synthesis_ref(gen, w, n), torch.float32)
    torch.cuda.synchronize()
    assert out.shape == (1, 9, 256, 256) and torch.isfinite(out).all()
    eo = float((out.double() - out64).abs().max() / out64.abs().max())
    eo32 = float((out32.double() - out64).abs().max() / out64.abs().max())
    print(f"synthesis output: max error {eo:.2e} of max |out| (torch FP32 {eo32:.2e})")
    assert eo <= max(1e-5, 2 * eo32)
    flips, total = _check_kinks(masks, pre64, "synthesis")
    print(f"synthesis: {flips} leaky-ReLU branch flips in {total} activations")
    _gradient_check(gw, gw64m, gw64, gw32, "dL/dW+")
    _gradient_check(gn, gn64m, gn64, gn32, "dL/dnoise")


@pytest.mark.parametrize("which", ["vgg_loss", "vgg_shaped"])
def test_vgg_stacks_and_input_gradient(pkg, which):
    """The two VGG stacks on the HIP convolutions (9 images of 256^2, as the
    MaterialGAN and feature-loss legs run them): features against the torch
    restatement in FP32 and FP64, dL/dimages against FP64 with the same ReLU
    branches (with the branch flips bounded as in the generator test)."""
    from svbrdf_amd import compose, materialgan as MG
    g = torch.Generator().manual_seed(9)
    ims = torch.rand((9, 3, 256, 256), generator=g).to(DEV)
    if which == "vgg_loss":
        net = MG.VGGLoss(DEV, np.array([1, 1, 4, 8]) / 14)
        fwd = lambda x: net.compute_feature_vector(net.normalize(x))   # noqa: E731
        ref = lambda x, **k: R.vgg_features_ref(net, x, **k)           # noqa: E731
    else:
        net = compose.VGGShapedFeatures(seed=4, device=DEV)
        fwd = lambda x: torch.cat([f.flatten() for f in net(x)])       # noqa: E731
        ref = lambda x, **k: torch.cat([f.flatten() for f in R.vgg_shaped_ref(net, x, **k)])  # noqa: E731
    probe = None

    def run(fn, dtype):
        nonlocal probe
        x = ims.detach().to(dtype).clone().requires_grad_(True)
        f = fn(x)
        if probe is None:
            probe = torch.randn(f.shape, generator=torch.Generator().manual_seed(10)).to(DEV)
        (f * probe.to(dtype)).sum().backward()
        return f.detach(), x.grad

    masks, remove = _masks_of(net.convs)
    f, gx = run(fwd, torch.float32)
    remove()
    pre64 = []
    f64, gx64 = run(lambda x: ref(x, record=pre64), torch.float64)
    _, gx64m = run(lambda x: ref(x, masks=masks), torch.float64)
    f32, gx32 = run(ref, torch.float32)
    torch.cuda.synchronize()
    ef = float((f.double() - f64).norm() / f64.norm())
    ef32 = float((f32.double() - f64).norm() / f64.norm())
    print(f"{which} features: {ef:.2e} (torch FP32 {ef32:.2e})")
    assert ef <= max(1e-5, 2 * ef32)
    flips, total = _check_kinks(masks, pre64, which)
    print(f"{which}: {flips} ReLU branch flips in {total} activations")
    _gradient_check(gx, gx64m, gx64, gx32, f"{which} dL/dimages")
