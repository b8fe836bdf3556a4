"""Pins the CPU oracle (oracle/) before anything is checked against it.

Each oracle function is checked against something other than itself:
* golden vectors written by the reference's own Microfacet class
  (tests/golden/make_golden.py): forward output and vector-Jacobian
  products in FP32 and FP64, and 5-step optimisation trajectories driven by
  torch.optim.Adam;
* closed-form known answers of the renderer;
* central finite differences of the FP64 oracle and torch.autograd.gradcheck
  of the eager restatement;
* torch.optim.Adam and torch's MSE loss themselves.
"""
import numpy as np
import pytest
import torch

from oracle import microfacet_np as M
from oracle import microfacet_torch as MT

FWD_CASES = ["renderjson_r8", "grid_r16", "edge_r8", "grid_r13"]


def load(golden_dir, name):
    z = np.load(golden_dir / f"fwd_vjp_{name}.npz")
    args = (z["tex"], int(z["res"]), float(z["size"]), z["cam"], z["light"], z["light_pow"])
    return z, args


@pytest.mark.parametrize("name", FWD_CASES)
def test_forward_fp64_matches_reference(golden_dir, name):
    z, args = load(golden_dir, name)
    out = M.render(*args, np.float64)
    np.testing.assert_allclose(out, z["out64"], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("name", FWD_CASES)
def test_forward_fp32_matches_reference_within_conditioning(golden_dir, name):
    """An independent FP32 evaluation differs from the reference's FP32 only
    by rounding; the bound is 1e-6 + 1e-5|ref| plus 8x the per-element FP32
    sensitivity of the ill-conditioned dot products (fp32_noise_scale)."""
    z, args = load(golden_dir, name)
    out = M.render(*args, np.float32)
    sig = M.fp32_noise_scale(*args)
    err = np.abs(out - z["out32"])
    assert np.all(err <= 1e-6 + 1e-5 * np.abs(z["out32"]) + 8 * sig)
    # and it is as close to the FP64 truth as the reference's own FP32 result
    e_ours = (np.abs(out - z["out64"]) / z["out64"]).max()
    e_ref = (np.abs(z["out32"] - z["out64"]) / z["out64"]).max()
    assert e_ours <= 2 * e_ref + 1e-6


@pytest.mark.parametrize("name", FWD_CASES)
def test_vjp_fp64_matches_reference_autograd(golden_dir, name):
    z, args = load(golden_dir, name)
    dtex, dP = M.render_vjp(*args, z["gout"], np.float64)
    scale = np.abs(z["dtex64"]).max(axis=(1, 2), keepdims=True)
    assert np.all(np.abs(dtex - z["dtex64"]) <= 1e-10 * scale + 1e-300)
    np.testing.assert_allclose(dP, z["dpow64"], rtol=1e-10)


@pytest.mark.parametrize("name", FWD_CASES)
def test_vjp_fp32_normwise(golden_dir, name):
    z, args = load(golden_dir, name)
    dtex, dP = M.render_vjp(*args, z["gout"], np.float32)
    for c in range(9):
        ref32, ref64 = z["dtex32"][c], z["dtex64"][c]
        n64 = np.linalg.norm(ref64)
        # no worse than the reference's own FP32 gradient (x2) or 1e-4 relative
        assert np.linalg.norm(dtex[c] - ref64) <= max(1e-4 * n64, 2 * np.linalg.norm(ref32 - ref64))
    np.testing.assert_allclose(dP, z["dpow64"], rtol=1e-4)


@pytest.mark.parametrize("name", FWD_CASES)
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_eager_restatement_matches_reference(golden_dir, name, dtype):
    z, args = load(golden_dir, name)
    tex, res, size, cam, light, pw = args
    sfx = "32" if dtype == torch.float32 else "64"
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(pw)]
    r = MT.MicrofacetEager(res, cam.shape[0], size, cl, "cpu", dtype)
    p = torch.tensor(pw, dtype=dtype, requires_grad=True)
    r.update_light(p)
    t = torch.tensor(tex, dtype=dtype, requires_grad=True)
    out = r.eval(t[None])
    (out * torch.tensor(z["gout"], dtype=dtype)).sum().backward()
    # same op sequence on the same library: identical to the last bit
    np.testing.assert_array_equal(out.detach().numpy(), z["out" + sfx])
    np.testing.assert_array_equal(t.grad.numpy(), z["dtex" + sfx])
    np.testing.assert_array_equal(p.grad.numpy(), z["dpow" + sfx])


# ---------------------------------------------------------------- closed forms
def _center_scene(z_height=20.0, res=5, size=4.0):
    cam = np.array([[0.0, 0.0, z_height]], np.float32)
    return cam, cam.copy(), res, size


def test_closed_form_flat_normal_under_light():
    """Flat normal, pixel directly below a co-located light: v = l = h = n,
    every cosine is 1, G = 1, D = a2/(pi a2^2 + eps),
    F = spe + (1 - spe) 2^(-5.55473 - 6.98316)."""
    cam, light, res, size = _center_scene()
    rng = np.random.default_rng(5)
    tex = np.zeros((9, res, res), np.float32)
    tex[0:3] = rng.uniform(-0.8, 0.8, (3, 1, 1))
    tex[5] = rng.uniform(-0.6, 0.6)
    tex[6:9] = rng.uniform(-0.9, -0.4, (3, 1, 1))
    P = np.array([300.0, 400.0, 500.0], np.float32)
    out = M.render(tex, res, size, cam, light, P, np.float64)
    c = res // 2
    t = tex[:, c, c].astype(np.float64)
    dif = ((t[0:3] + 1) / 2) ** 2.2
    spe = ((t[6:9] + 1) / 2) ** 2.2
    rgh = ((t[5] + 1) / 2) ** 2.2
    a2 = rgh ** 4
    D = a2 / (np.pi * a2 * a2 + 1e-6)
    F = spe + (1 - spe) * 2.0 ** (-5.55473 - 6.98316)
    img = P * (dif / np.pi * (1 - spe) + D * F / (4 + 1e-6)) / float(cam[0, 2]) ** 2
    np.testing.assert_allclose(out[0, :, c, c], np.clip(img, 1e-6, 1) ** (1 / 2.2), rtol=1e-12)


def test_closed_form_black_textures():
    """dif = spe = rgh = 0 gives img = 0, so every pixel is eps^(1/2.2)."""
    cam = np.array([[1, 2, 20], [-3, 0, 15]], np.float32)
    tex = np.full((9, 6, 6), -1.0, np.float32)
    tex[3:5] = np.random.default_rng(0).uniform(-0.3, 0.3, (2, 6, 6))
    out = M.render(tex, 6, 10.0, cam, cam, np.full(3, 1500, np.float32), np.float64)
    np.testing.assert_allclose(out, 1e-6 ** (1 / 2.2), rtol=1e-14)
    assert abs(1e-6 ** (1 / 2.2) - 1.8738e-3) < 1e-7


def test_closed_form_saturation_has_zero_gradient():
    cam = np.array([[0, 0, 20]], np.float32)
    tex = np.zeros((9, 4, 4), np.float32)
    P = np.full(3, 1e9, np.float32)
    out = M.render(tex, 4, 4.0, cam, cam, P, np.float64)
    np.testing.assert_array_equal(out, 1.0)
    dtex, dP = M.render_vjp(tex, 4, 4.0, cam, cam, P, np.ones_like(out), np.float64)
    np.testing.assert_array_equal(dtex, 0.0)
    np.testing.assert_array_equal(dP, 0.0)


def test_closed_form_back_facing_light():
    cam = np.array([[0, 0, 20]], np.float32)
    light = np.array([[0, 0, -20]], np.float32)
    tex = np.zeros((9, 4, 4), np.float32)
    out = M.render(tex, 4, 4.0, cam, light, np.full(3, 1500, np.float32), np.float64)
    np.testing.assert_allclose(out, 1e-6 ** (1 / 2.2), rtol=1e-14)


def test_mirror_symmetry():
    """Mirroring the scene in x (texture columns, nx sign, camera x) mirrors
    the image; catches sign/index mistakes in the pixel grid and normals."""
    rng = np.random.default_rng(9)
    res = 8
    tex = rng.uniform(-0.9, 0.9, (9, res, res)).astype(np.float32)
    tex[3:5] *= 0.5
    cam = np.array([[-3, 2, 18], [4, -1, 25]], np.float32)
    out = M.render(tex, res, 7.0, cam, cam, np.full(3, 1500, np.float32), np.float64)
    tm = tex[:, :, ::-1].copy()
    tm[3] = -tm[3]
    cm = cam * np.array([-1, 1, 1], np.float32)
    outm = M.render(tm, res, 7.0, cm, cm, np.full(3, 1500, np.float32), np.float64)
    np.testing.assert_allclose(outm, out[:, :, :, ::-1], rtol=1e-12)


# ------------------------------------------------------- finite differences
@pytest.mark.parametrize("mode", [M.LOSS_L2, M.LOSS_LOG_L2])
def test_gradient_matches_central_differences(mode):
    rng = np.random.default_rng(3)
    res = 6
    tex = rng.uniform(-0.9, 0.9, (9, res, res))
    tex[3:5] = rng.uniform(-0.4, 0.4, (2, res, res))
    cam = np.array([[x, y, 12.0] for y in (3.0, -3.0) for x in (-3.0, 3.0)])
    P = np.array([900.0, 1100.0, 1300.0])
    target = M.render(np.clip(tex + rng.normal(0, 0.1, tex.shape), -1, 1), res, 8.0, cam, cam, P, np.float64)

    def L(t):
        return M.loss(M.render(np.clip(t, -1, 1), res, 8.0, cam, cam, P, np.float64), target, mode, np.float64)

    out = M.render(tex, res, 8.0, cam, cam, P, np.float64)
    g, _ = M.render_vjp(tex, res, 8.0, cam, cam, P, M.loss_grad(out, target, mode, np.float64), np.float64)
    h = 1e-5
    L0 = L(tex)
    # FD error: truncation O(h^2) plus roundoff ~ 8 * 2^-52 * |L| / h
    roundoff = 8 * np.finfo(np.float64).eps * abs(L0) / h
    picks = [(c, i, j) for c in range(9) for (i, j) in [(1, 2), (4, 3), (5, 0)]]
    for c, i, j in picks:
        tp, tm = tex.copy(), tex.copy()
        tp[c, i, j] += h
        tm[c, i, j] -= h
        fd = (L(tp) - L(tm)) / (2 * h)
        assert abs(fd - g[c, i, j]) <= 1e-5 * abs(fd) + roundoff, (c, i, j, fd, g[c, i, j])


def test_gradcheck_eager_restatement():
    torch.manual_seed(0)
    res = 3
    cam = torch.tensor([[1.0, -1.0, 6.0], [-2.0, 0.5, 5.0]], dtype=torch.float64)
    light = torch.tensor([[0.5, 1.0, 7.0], [-2.0, 0.5, 5.0]], dtype=torch.float64)
    pw = torch.tensor([40.0, 50.0, 60.0], dtype=torch.float64, requires_grad=True)
    r = MT.MicrofacetEager(res, 2, 3.0, [cam, light, pw.detach()], "cpu", torch.float64)
    tex = (torch.rand(1, 9, res, res, dtype=torch.float64) * 1.6 - 0.8)
    tex[:, 3:5] *= 0.4
    tex.requires_grad_(True)

    def f(t, p):
        r.update_light(p)
        return r.eval(t)

    assert torch.autograd.gradcheck(f, (tex, pw), eps=1e-6, atol=1e-7, rtol=1e-5)


# ------------------------------------------------------------ losses / Adam
def test_losses_match_torch():
    rng = np.random.default_rng(1)
    out = rng.uniform(0.002, 1, (3, 3, 4, 4)).astype(np.float32)
    tgt = rng.uniform(0.0, 1, (3, 3, 4, 4)).astype(np.float32)
    tgt[0, 0, 0, 0] = 0.0           # below the renderer's floor: clamped
    for mode in (M.LOSS_L2, M.LOSS_LOG_L2):
        o = torch.tensor(out, dtype=torch.float64, requires_grad=True)
        Lt = MT.image_loss(o, torch.tensor(tgt, dtype=torch.float64), mode)
        Lt.backward()
        np.testing.assert_allclose(M.loss(out, tgt, mode, np.float64), Lt.item(), rtol=1e-12)
        np.testing.assert_allclose(M.loss_grad(out, tgt, mode, np.float64), o.grad.numpy(), rtol=1e-12)


def test_adam_matches_torch_optim():
    rng = np.random.default_rng(2)
    p0 = rng.normal(size=(9, 5, 5)).astype(np.float32)
    p = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([p], lr=0.01, betas=(0.9, 0.999), foreach=False)
    q, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for step in range(1, 8):
        g = (rng.normal(size=p0.shape) * 10.0 ** rng.uniform(-10, -6, p0.shape)).astype(np.float32)
        p.grad = torch.tensor(g)
        opt.step()
        q, m, v = M.adam_step(q, g, m, v, step, 0.01)
        st = opt.state[p]
        # m, v: same rounding as torch's CPU kernels (fused multiply-adds)
        np.testing.assert_array_equal(m, st["exp_avg"].numpy())
        np.testing.assert_array_equal(v, st["exp_avg_sq"].numpy())
        # params: torch's vectorised sqrt is not always correctly rounded, so
        # the denominators may differ by an ulp
        np.testing.assert_allclose(q, p.detach().numpy(), rtol=0, atol=2e-7)


def test_adam_first_step_closed_form():
    g = np.array([1e-3, -2e-7, 5e-9, 0.0], np.float64)
    p, _, _ = M.adam_step(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, 0.01, dtype=np.float64)
    np.testing.assert_allclose(p, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)


# --------------------------------------------------- optimisation trajectories
@pytest.mark.parametrize("loss", ["l2", "logl2"])
@pytest.mark.parametrize("light", [False, True])
def test_fused_step_oracle_matches_reference_loop(golden_dir, loss, light):
    """The oracle's clamp/render/loss/backward/Adam step against 5 iterations
    of the reference renderer driven by torch.optim.Adam."""
    z = np.load(golden_dir / f"optim_{loss}_r8.npz")
    mode = M.LOSS_L2 if loss == "l2" else M.LOSS_LOG_L2
    key = "light" if light else "tex"
    res, size = int(z["res"]), float(z["size"])
    t = z["init"].copy()
    m = np.zeros_like(t)
    v = np.zeros_like(t)
    P = z["light_pow0"].copy()
    ls = (np.zeros(3, np.float32), np.zeros(3, np.float32))
    for k in range(5):
        if light:
            t, m, v, L, (P, lm, lv) = M.fused_step(t, m, v, z["target"], k + 1, res, size, z["cam"], z["light"], P,
                                                   mode, float(z["lr"]), light_state=ls)
            ls = (lm, lv)
        else:
            t, m, v, L = M.fused_step(t, m, v, z["target"], k + 1, res, size, z["cam"], z["light"], P, mode,
                                      float(z["lr"]))
        np.testing.assert_allclose(L, z[f"{key}_loss"][k], rtol=2e-5)
        ref = z[f"{key}_textures"][k]
        # Adam's first steps move parameters by ~lr * g/(|g|+eps): entries whose
        # gradient is ~eps can differ by up to lr; all others agree to 1e-5
        close = np.abs(t - ref) <= 1e-5
        assert close.mean() >= 0.995, (k, close.mean())
        assert np.abs(t - ref).max() <= 2 * float(z["lr"]) * (k + 1)
        if light:
            np.testing.assert_allclose(P, z[f"{key}_light_pow"][k], rtol=1e-6)


# ------------------------------------------ alternate terms (N4): Beckmann, Schlick
def _single_pixel_terms(theta, rgh, spe, vh=None, ndf="ggx", fresnel="sg"):
    """_terms for one pixel with n = +z and a half vector at each angle theta
    (one 'view' per angle, v = l = h)."""
    V = len(theta)
    h = np.stack([np.sin(theta), np.zeros(V), np.cos(theta)], 1)[:, :, None, None]
    geo = dict(v=h, l=h, h=h, vh=np.ones((V, 1, 1, 1)) if vh is None else vh.reshape(V, 1, 1, 1),
               d2=np.ones((V, 1, 1, 1)))
    mp = dict(n=np.array([0.0, 0.0, 1.0])[:, None, None], dif=np.zeros((3, 1, 1)),
              spe=np.full((3, 1, 1), spe), rgh=np.full((1, 1, 1), rgh))
    return M._terms(mp, geo, np.ones(3), np.float64, ndf=ndf, fresnel=fresnel)


@pytest.mark.parametrize("ndf", ["ggx", "beckmann"])
@pytest.mark.parametrize("rgh", [0.4, 0.6, 0.8])
def test_ndf_projected_area_normalisation(ndf, rgh):
    # int D(h) (n.h) dw over the hemisphere = 1 for both microfacet distributions
    theta = np.linspace(0, np.pi / 2, 40001)[:-1]
    D = _single_pixel_terms(theta, rgh, 0.04, ndf=ndf)["D"][:, 0, 0, 0]
    integrand = D * np.cos(theta) * np.sin(theta) * 2 * np.pi
    total = np.sum((integrand[1:] + integrand[:-1]) / 2 * np.diff(theta))
    assert abs(total - 1) < 2e-3, total


def test_beckmann_and_schlick_closed_forms():
    alpha = 0.5 ** 2
    tm = _single_pixel_terms(np.array([0.0, np.arctan(0.3)]), 0.5, 0.04, ndf="beckmann")
    # theta = 0: D = 1/(pi a2); tan^2 = 0.09: D = exp(-0.09/a2)/(pi a2 cos^4)
    np.testing.assert_allclose(tm["D"][0, 0, 0, 0], 1 / (np.pi * alpha ** 2), rtol=1e-12)
    c2 = np.cos(np.arctan(0.3)) ** 2
    np.testing.assert_allclose(tm["D"][1, 0, 0, 0], np.exp(-0.09 / alpha ** 2) / (np.pi * alpha ** 2 * c2 ** 2),
                               rtol=1e-12)
    vh = np.array([1.0, 0.0, 0.5])
    tm = _single_pixel_terms(np.zeros(3), 0.5, 0.04, vh=vh, fresnel="schlick")
    np.testing.assert_allclose(tm["F"][:, 0, 0, 0], [0.04, 1.0, 0.04 + 0.96 / 32], rtol=1e-12)


@pytest.mark.parametrize("ndf,fresnel", [("beckmann", "sg"), ("ggx", "schlick"), ("beckmann", "schlick")])
def test_variant_gradients_match_central_differences(ndf, fresnel):
    rng = np.random.default_rng(13)
    res = 5
    tex = rng.uniform(-0.8, 0.8, (9, res, res))
    tex[3:5] = rng.uniform(-0.4, 0.4, (2, res, res))
    cam = np.array([[-2.0, 1.0, 10.0], [3.0, -2.0, 11.0]])
    light = np.array([[1.0, 2.0, 9.0], [-3.0, 1.5, 12.0]])
    P = np.array([400.0, 500.0, 600.0])
    gout = rng.normal(0, 1, (2, 3, res, res))

    def L(t):
        return float(np.sum(M.render(t, res, 6.0, cam, light, P, np.float64, ndf=ndf, fresnel=fresnel) * gout))

    g, _ = M.render_vjp(tex, res, 6.0, cam, light, P, gout, np.float64, ndf=ndf, fresnel=fresnel)
    h = 1e-6
    for c, i, j in [(c, i, j) for c in range(9) for (i, j) in [(0, 1), (3, 4), (2, 2)]]:
        tp, tm = tex.copy(), tex.copy()
        tp[c, i, j] += h
        tm[c, i, j] -= h
        fd = (L(tp) - L(tm)) / (2 * h)
        assert abs(fd - g[c, i, j]) <= 1e-6 * max(1.0, abs(fd)), (ndf, fresnel, c, i, j, fd, g[c, i, j])
