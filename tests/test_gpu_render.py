"""Forward render and adjoint on the GPU (C ABI through the drop-in
Microfacet) against the reference's own outputs (golden vectors), the
eager restatement (bit-identical to the reference in FP32) and the FP64
oracle.  Criteria and tolerances: tests/parity.py."""
import numpy as np
import pytest
import torch

from oracle import microfacet_np as M
from oracle import microfacet_torch as MT
from tests import parity

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
GOLDEN = ["renderjson_r8", "grid_r16", "edge_r8", "grid_r13"]


def load_golden(golden_dir, name):
    z = np.load(golden_dir / f"fwd_vjp_{name}.npz")
    return z, (z["tex"], int(z["res"]), float(z["size"]), z["cam"], z["light"], z["light_pow"])


def scene(pkg, res, V, seed=0, coloc=True):
    tex = pkg.synthetic.ground_truth_textures(res, seed=seed)
    rng = np.random.default_rng(seed + 100)
    cam, light = pkg.synthetic.flash_grid(jitter=3.0 if V != 9 else 0.0, rng=rng, n_views=V)
    if not coloc:
        light = (light + rng.uniform(-2, 2, light.shape)).astype(np.float32)
    return tex, res, pkg.synthetic.PLANE_SIZE_CM, cam, light, pkg.synthetic.light_power()


def gpu_render(pkg, tex, res, size, cam, light, P, gout=None):
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    r = pkg.Microfacet(res, cam.shape[0], size, cl, DEV)
    t = torch.tensor(tex, device=DEV)[None].requires_grad_(gout is not None)
    p = torch.tensor(P, device=DEV).requires_grad_(gout is not None)
    r.update_light(p)
    out = r.eval(t)
    if gout is None:
        torch.cuda.synchronize()
        return out.detach().cpu().numpy()
    (out * torch.tensor(gout, device=DEV)).sum().backward()
    torch.cuda.synchronize()
    return out.detach().cpu().numpy(), t.grad[0].cpu().numpy(), p.grad.cpu().numpy()


def eager_vjp(tex, res, size, cam, light, P, gout, dtype):
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    r = MT.MicrofacetEager(res, cam.shape[0], size, cl, "cpu", dtype)
    p = torch.tensor(P, dtype=dtype, requires_grad=True)
    r.update_light(p)
    t = torch.tensor(tex, dtype=dtype, requires_grad=True)
    out = r.eval(t[None])
    (out * torch.tensor(gout, dtype=dtype)).sum().backward()
    return out.detach().numpy(), t.grad.numpy(), p.grad.numpy()


@pytest.mark.parametrize("name", GOLDEN)
def test_forward_matches_reference_golden(pkg, golden_dir, name):
    z, args = load_golden(golden_dir, name)
    out = gpu_render(pkg, *args)
    parity.check_forward(out, z["out32"], z["out64"], M.fp32_noise_scale(*args), name)


@pytest.mark.parametrize("name", GOLDEN)
def test_adjoint_matches_reference_autograd_golden(pkg, golden_dir, name):
    z, args = load_golden(golden_dir, name)
    out, dtex, dP = gpu_render(pkg, *args, gout=z["gout"])
    parity.check_gradient(dtex, z["dtex32"], z["dtex64"], name)
    np.testing.assert_allclose(dP, z["dpow64"], rtol=1e-4)


@pytest.mark.parametrize("res,V,coloc", [(1, 1, True), (2, 3, False), (13, 9, True), (64, 9, True),
                                         (64, 16, False), (100, 4, True), (256, 9, True),
                                         (5, 1024, True), (3, 1024, False)])
def test_forward_and_adjoint_synthetic(pkg, res, V, coloc):
    args = scene(pkg, res, V, seed=res + V, coloc=coloc)
    gout = np.random.default_rng(res).standard_normal((V, 3, res, res)).astype(np.float32)
    out, dtex, dP = gpu_render(pkg, *args, gout=gout)
    o32, g32, p32 = eager_vjp(*args, gout, torch.float32)
    o64, g64, p64 = eager_vjp(*args, gout, torch.float64)
    parity.check_forward(out, o32, o64, M.fp32_noise_scale(*args), f"R{res}V{V}")
    parity.check_gradient(dtex, g32, g64, f"R{res}V{V}")
    assert np.linalg.norm(dP - p64) <= max(1e-4 * np.linalg.norm(p64), 2 * np.linalg.norm(p32 - p64))


@pytest.mark.parametrize("e", [0, 63, 64, 89])
def test_one_offset_light_coordinate_takes_the_general_path(pkg, e):
    """Every light at its camera except coordinate e of the 3 V flattened
    positions (V = 30, so the co-location test spans two 64-coordinate
    chunks): the kernels must shade with the general geometry.  The
    co-located forms would light view e // 3 from its camera, 1.5 cm away."""
    res, V = 12, 30
    tex, res, size, cam, light, P = scene(pkg, res, V, seed=7)
    light = light.copy()
    light.reshape(-1)[e] += 1.5
    gout = np.random.default_rng(e).standard_normal((V, 3, res, res)).astype(np.float32)
    out, dtex, dP = gpu_render(pkg, tex, res, size, cam, light, P, gout=gout)
    o32, g32, _ = eager_vjp(tex, res, size, cam, light, P, gout, torch.float32)
    o64, g64, _ = eager_vjp(tex, res, size, cam, light, P, gout, torch.float64)
    parity.check_forward(out, o32, o64, M.fp32_noise_scale(tex, res, size, cam, light, P), f"offset e={e}")
    parity.check_gradient(dtex, g32, g64, f"offset e={e}")


def test_forward_1024_band_and_properties(pkg):
    """Forward render at 1024^2 x 9: elementwise parity on a band of rows,
    range and mirror symmetry on the whole image."""
    res, V = 1024, 9
    tex, _, size, cam, light, P = scene(pkg, res, V, seed=0)
    out = gpu_render(pkg, tex, res, size, cam, light, P)
    assert out.shape == (V, 3, res, res)
    assert np.isfinite(out).all()
    assert out.min() >= np.float32(1e-6) ** np.float32(1 / 2.2) * (1 - 1e-6) and out.max() <= 1.0
    rows = slice(496, 528)
    o32 = M.render(tex, res, size, cam, light, P, np.float32, rows=rows)
    o64 = M.render(tex, res, size, cam, light, P, np.float64, rows=rows)
    sig = M.fp32_noise_scale(tex, res, size, cam, light, P, rows=rows)
    parity.check_forward(out[:, :, rows], o32, o64, sig, "1024 band")
    # mirrored scene (columns reversed, nx negated, camera x negated) -> mirrored image, exactly
    tm = tex[:, :, ::-1].copy()
    tm[3] = -tm[3]
    cm = (cam * np.array([-1, 1, 1], np.float32)).astype(np.float32)
    outm = gpu_render(pkg, tm, res, size, cm, cm, P)
    np.testing.assert_array_equal(outm, out[:, :, :, ::-1])


def test_adjoint_1024_band(pkg):
    res, V = 1024, 9
    tex, _, size, cam, light, P = scene(pkg, res, V, seed=1)
    rng = np.random.default_rng(7)
    gout = rng.standard_normal((V, 3, res, res)).astype(np.float32)
    _, dtex, _ = gpu_render(pkg, tex, res, size, cam, light, P, gout=gout)
    rows = slice(200, 216)
    g64, _ = M.render_vjp(tex, res, size, cam, light, P, gout[:, :, rows], np.float64, rows=rows)
    g32, _ = M.render_vjp(tex, res, size, cam, light, P, gout[:, :, rows], np.float32, rows=rows)
    parity.check_gradient(dtex[:, rows], g32, g64, "1024 band")


def test_adjoint_light_gradient_1024(pkg):
    """dL/dP at 1024^2 x 9 views through all 16,384 one-wave blocks' partials
    and the fixed-order reduction.  dL/dimages is zero outside the first and
    last 64 rows, so the FP64 oracle needs only those rows; every block still
    writes a partial (zero between the bands)."""
    res, V = 1024, 9
    tex, _, size, cam, light, P = scene(pkg, res, V, seed=1)
    rng = np.random.default_rng(11)
    gout = np.zeros((V, 3, res, res), np.float32)
    bands = (slice(0, 64), slice(res - 64, res))
    for b in bands:
        gout[:, :, b] = rng.standard_normal((V, 3, 64, res)).astype(np.float32)
    _, dtex, dP = gpu_render(pkg, tex, res, size, cam, light, P, gout=gout)
    dP64 = np.zeros(3)
    for b in bands:
        g64, p64 = M.render_vjp(tex, res, size, cam, light, P, gout[:, :, b], np.float64, rows=b)
        g32, _ = M.render_vjp(tex, res, size, cam, light, P, gout[:, :, b], np.float32, rows=b)
        parity.check_gradient(dtex[:, b], g32, g64, f"1024 rows {b.start}-{b.stop}")
        dP64 += p64
    assert np.all(dtex[:, 64:res - 64] == 0.0)
    np.testing.assert_allclose(dP, dP64, rtol=1e-4)


def test_eval_is_deterministic_and_stream_ordered(pkg):
    args = scene(pkg, 128, 9, seed=3)
    a = gpu_render(pkg, *args)
    b = gpu_render(pkg, *args)
    np.testing.assert_array_equal(a, b)


def test_nan_propagates_like_reference(pkg):
    """A diffuse value below -1 makes ((t+1)/2)^2.2 NaN in the reference."""
    tex, res, size, cam, light, P = scene(pkg, 8, 2, seed=5)
    tex = tex.copy()
    tex[0, 3, 4] = -1.5
    out = gpu_render(pkg, tex, res, size, cam, light, P)
    ref = M.render(tex, res, size, cam, light, P, np.float64)
    np.testing.assert_array_equal(np.isnan(out), np.isnan(ref))
    assert np.isnan(out).any()
