"""Fused optimisation step (clamp, render, loss, adjoint, Adam in one kernel)
against the reference loop: the golden 5-step trajectories written with the
reference renderer and torch.optim.Adam, the eager restatement with
torch.optim.Adam, the FP64 oracle and FP64 central finite differences.

The gradient the kernel used is recovered exactly from the first step:
Adam's first moment after one step is m1 = (1 - beta1) g."""
import numpy as np
import pytest
import torch

from oracle import microfacet_np as M
from oracle import microfacet_torch as MT
from tests import parity
from tests.test_gpu_render import scene

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def make_problem(pkg, tex0, target, res, size, cam, light, P, loss, lr=0.01, optim_light=False, history=64, **kw):
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    r = pkg.Microfacet(res, cam.shape[0], size, cl, DEV)
    tex = torch.tensor(tex0, device=DEV).reshape(1, 9, res, res).contiguous()
    tg = torch.tensor(target, device=DEV)
    return pkg.FusedStep(r, tex, tg, lr, loss=loss, optim_light=optim_light, history=history, **kw), tex


def eager_loss_grad(tex, target, res, size, cam, light, P, mode, dtype, device="cpu", kinks=False):
    """Loss and torch-autograd gradient of the eager restatement of the
    reference renderer (the oracle; on the GPU for the large shapes); with
    kinks=True also the pixels at a kink of the output clamp
    (parity.clamp_kinks of the pre-clamp image)."""
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    r = MT.MicrofacetEager(res, cam.shape[0], size, cl, device, dtype)
    t = torch.as_tensor(tex).to(device=device, dtype=dtype).detach().requires_grad_(True)
    tg = torch.as_tensor(target).to(device=device, dtype=dtype)
    L = MT.image_loss(r.eval(t[None].clamp(-1, 1)), tg, mode)
    L.backward()
    g = t.grad.cpu().numpy()
    mask = None
    if kinks:
        with torch.no_grad():
            mask = parity.clamp_kinks(r.radiance(t.detach()[None].clamp(-1, 1)).cpu().numpy())
    del r, t, tg
    return (L.item(), g, mask) if kinks else (L.item(), g)


@pytest.mark.parametrize("loss,mode", [("l2", M.LOSS_L2), ("log_l2", M.LOSS_LOG_L2)])
@pytest.mark.parametrize("res,V,coloc", [(64, 9, True), (33, 5, False), (256, 9, True), (1, 1, True),
                                         (4, 1024, True), (3, 1024, False)])
def test_first_step_gradient_and_loss(pkg, loss, mode, res, V, coloc):
    gt, _, size, cam, light, P = scene(pkg, res, V, seed=11, coloc=coloc)
    init = pkg.synthetic.init_textures(gt, seed=1)
    target = M.render(gt, res, size, cam, light, P, np.float32)
    fs, tex = make_problem(pkg, init, target, res, size, cam, light, P, loss)
    fs.step()
    torch.cuda.synchronize()
    g_gpu = fs.exp_avg[0].cpu().numpy().astype(np.float64) / np.float64(np.float32(0.1))
    L32, g32 = eager_loss_grad(init, target, res, size, cam, light, P, mode, torch.float32)
    L64, g64 = eager_loss_grad(init, target, res, size, cam, light, P, mode, torch.float64)
    parity.check_gradient(g_gpu, g32, g64, f"step1 {loss} R{res}")
    assert abs(fs.loss() - L64) <= 1e-5 * abs(L64)
    # Adam's first update is -lr * g / (|g| + eps) (bias corrections cancel)
    st = fs.read_state()
    assert st["step"] == 1.0
    p_exp = init - 0.01 * g64 / (np.abs(g64) + 1e-8)
    assert np.mean(np.abs(tex[0].cpu().numpy() - p_exp) <= 1e-5) >= 0.999


@pytest.mark.parametrize("e", [0, 89])
def test_one_offset_light_coordinate_step(pkg, e):
    """The fused step's co-location test over two 64-coordinate chunks
    (V = 30): one light coordinate off its camera selects the general path."""
    res, V = 12, 30
    gt, _, size, cam, light, P = scene(pkg, res, V, seed=5)
    light = light.copy()
    light.reshape(-1)[e] += 1.5
    init = pkg.synthetic.init_textures(gt, seed=1)
    target = M.render(gt, res, size, cam, light, P, np.float32)
    fs, _ = make_problem(pkg, init, target, res, size, cam, light, P, "log_l2")
    fs.step()
    torch.cuda.synchronize()
    g_gpu = fs.exp_avg[0].cpu().numpy().astype(np.float64) / np.float64(np.float32(0.1))
    _, g32 = eager_loss_grad(init, target, res, size, cam, light, P, M.LOSS_LOG_L2, torch.float32)
    L64, g64 = eager_loss_grad(init, target, res, size, cam, light, P, M.LOSS_LOG_L2, torch.float64)
    parity.check_gradient(g_gpu, g32, g64, f"offset step e={e}")
    assert abs(fs.loss() - L64) <= 1e-5 * abs(L64)


@pytest.mark.parametrize("loss", ["l2", "logl2"])
@pytest.mark.parametrize("light", [False, True])
def test_matches_reference_loop_golden(pkg, golden_dir, loss, light):
    """5 free-running steps at R=8 against the reference renderer +
    torch.optim.Adam.  The two trajectories are compared as wholes: losses to
    1e-4, 99 % of the maps to 1e-5, and no value further apart than two
    Adam trajectories can move in k steps (2 lr k: Adam's step is at most ~lr
    per iteration here).  Individual values drift apart where a gradient is
    ill-conditioned (its FP32 error is 1e-3..1e-1 relative in any FP32
    implementation, the reference's included) and Adam carries the
    difference on; test_golden_trajectory_step_by_step checks every value of
    every step from the reference's own states, with no allowance."""
    z = np.load(golden_dir / f"optim_{loss}_r8.npz")
    key = "light" if light else "tex"
    res, size = int(z["res"]), float(z["size"])
    fs, tex = make_problem(pkg, z["init"], z["target"], res, size, z["cam"], z["light"], z["light_pow0"],
                           loss, lr=float(z["lr"]), optim_light=light)
    for k in range(5):
        fs.step()
        torch.cuda.synchronize()
        got = tex[0].cpu().numpy()
        ref = z[f"{key}_textures"][k]
        close = np.abs(got - ref) <= 1e-5
        assert close.mean() >= 0.99, (k, close.mean())
        assert np.abs(got - ref).max() <= 2 * float(z["lr"]) * (k + 1)
        if light:
            np.testing.assert_allclose(fs.light_pow.cpu().numpy(), z[f"{key}_light_pow"][k], rtol=1e-5)
    np.testing.assert_allclose(fs.losses(), z[f"{key}_loss"], rtol=1e-4)


def test_config0_twenty_steps_against_eager_torch_adam(pkg):
    """256^2, 9 views, 20 free-running Adam iterations at lr 0.01, log-space
    L2; the reference loop is the eager restatement driven by
    torch.optim.Adam.  Compared as whole trajectories (losses to 1e-4,
    99.9 % of the maps to 1e-4, nothing further apart than the 2 x 20 lr two
    Adam trajectories can open); the elementwise check of every step is
    test_config0_trajectory_step_by_step."""
    res = 256
    gt, _, size, cam, light, P = scene(pkg, res, 9, seed=0)
    init = pkg.synthetic.init_textures(gt, seed=1)
    target = M.render(gt, res, size, cam, light, P, np.float32)
    fs, tex = make_problem(pkg, init, target, res, size, cam, light, P, "log_l2")
    for _ in range(20):
        fs.step()
    torch.cuda.synchronize()
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    r = MT.MicrofacetEager(res, 9, size, cl, "cpu", torch.float32)
    t_ref, _, _, losses = MT.optimise(r, torch.tensor(init)[None], torch.tensor(target), 20, 0.01,
                                      MT.LOSS_LOG_L2, foreach=False)
    got = tex[0].cpu().numpy()
    ref = t_ref[0].numpy()
    err = np.abs(got - ref)
    assert np.mean(err <= 1e-4) >= 0.999, np.mean(err <= 1e-4)
    assert err.max() <= 0.4
    np.testing.assert_allclose(fs.losses(), [l.item() for l in losses], rtol=1e-4)
    assert fs.losses()[-1] < fs.losses()[0]


def test_graph_replay_matches_eager_steps(pkg):
    res = 128
    gt, _, size, cam, light, P = scene(pkg, res, 9, seed=4)
    init = pkg.synthetic.init_textures(gt, seed=1)
    target = M.render(gt, res, size, cam, light, P, np.float32)
    fa, ta = make_problem(pkg, init, target, res, size, cam, light, P, "log_l2")
    fb, tb = make_problem(pkg, init, target, res, size, cam, light, P, "log_l2")
    for _ in range(3):
        fa.step()
    fb.capture()
    for k in range(3):
        fb.replay()
        # eager work and a host round trip between replays
        _ = (tb.sum() * 0).item()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ta.cpu().numpy(), tb.cpu().numpy())
    np.testing.assert_array_equal(fa.exp_avg_sq.cpu().numpy(), fb.exp_avg_sq.cpu().numpy())
    np.testing.assert_array_equal(fa.losses(), fb.losses())
    assert fb.read_state()["step"] == 3.0


@pytest.mark.parametrize("res,V,coloc,light", [(33, 5, False, False), (128, 9, True, False), (64, 9, True, True)])
def test_plane_layouts_are_bit_identical(pkg, res, V, coloc, light):
    """Moments and targets at the padded plane stride, dense (0) or an odd
    stride, and the maps in the caller's tensor or in the step's own padded
    copy: the same arithmetic on the same values."""
    gt, _, size, cam, light_pos, P = scene(pkg, res, V, seed=6, coloc=coloc)
    init = pkg.synthetic.init_textures(gt, seed=2)
    target = M.render(gt, res, size, cam, light_pos, P, np.float32)
    runs = []
    for ps, own in ((None, False), (0, False), (res * res + 37, False), (None, True), (0, True)):
        fs, tex = make_problem(pkg, init, target, res, size, cam, light_pos, P, "log_l2", optim_light=light,
                               plane_stride=ps, own_textures=own)
        assert fs.plane_stride == (pkg._lib.query("svbrdf_plane_stride", res) if ps is None else ps)
        assert fs.texture_stride == (fs.plane_stride if own else 0)
        for _ in range(4):
            fs.step()
        torch.cuda.synchronize()
        if own:
            # the caller's tensor is untouched until write_back
            assert np.array_equal(tex.cpu().numpy()[0], init)
            fs.write_back()
        runs.append((tex.cpu().numpy(), fs.textures.cpu().numpy(), fs.exp_avg.cpu().numpy(),
                     fs.exp_avg_sq.cpu().numpy(), fs.losses(), fs.light_pow.cpu().numpy()))
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            np.testing.assert_array_equal(a, b)


def test_deterministic(pkg):
    res = 200
    gt, _, size, cam, light, P = scene(pkg, res, 9, seed=6)
    init = pkg.synthetic.init_textures(gt, seed=1)
    target = M.render(gt, res, size, cam, light, P, np.float32)
    outs = []
    for _ in range(2):
        fs, tex = make_problem(pkg, init, target, res, size, cam, light, P, "l2", optim_light=True)
        for _ in range(4):
            fs.step()
        torch.cuda.synchronize()
        outs.append((tex.cpu().numpy(), fs.losses(), fs.light_pow.cpu().numpy()))
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a, b)


def raw_state(pkg, fs):
    """Device step state without the flush that read_state() performs."""
    dt = pkg.optim._STATE_DTYPE
    return fs.state.detach().cpu().numpy().view(np.uint8)[:dt.itemsize].view(dt)[0]


@pytest.mark.parametrize("res", [16, 200, 1024])
def test_single_launch_records_losses_one_step_late_and_flush_completes(pkg, res):
    """Without light optimisation a step is one launch: step t's partial sums
    are reduced by the launch of step t + 1 (block 0) or by the flush.  Both
    use one fixed-order reduction, so flushing after every step gives the
    same losses bit for bit, and the arrival counters are back at 0 after
    every launch."""
    gt, _, size, cam, light, P = scene(pkg, res, 9, seed=8)
    init = pkg.synthetic.init_textures(gt, seed=2)
    target = M.render(gt, res, size, cam, light, P, np.float32)
    fa, ta = make_problem(pkg, init, target, res, size, cam, light, P, "log_l2", history=8)
    fb, tb = make_problem(pkg, init, target, res, size, cam, light, P, "log_l2", history=8)
    for _ in range(3):
        fa.step()
        fb.step()
        fb.flush()
    torch.cuda.synchronize()
    hist = fa.history.cpu().numpy()
    assert np.all(hist[:2] > 0) and hist[2] == 0.0
    st = raw_state(pkg, fa)
    assert st["step"] == 3.0 and st["pending_step"] == 3 and st["loss"] == hist[1]
    # all 65 arrival counters (one per 128-byte line) are back at zero
    assert int(fa.workspace.view(torch.int32)[:65 * 32].abs().sum().item()) == 0
    assert raw_state(pkg, fb)["pending_step"] == 0
    la = fa.losses()
    np.testing.assert_array_equal(la, fb.losses())
    assert la[2] > 0 and fa.loss() == la[2]
    assert raw_state(pkg, fa)["pending_step"] == 0
    # a second flush changes nothing
    fa.flush()
    np.testing.assert_array_equal(fa.losses(), la)
    np.testing.assert_array_equal(ta.cpu().numpy(), tb.cpu().numpy())


def test_state_dict_resume_in_a_fresh_workspace(pkg):
    res = 96
    gt, _, size, cam, light, P = scene(pkg, res, 9, seed=9)
    init = pkg.synthetic.init_textures(gt, seed=3)
    target = M.render(gt, res, size, cam, light, P, np.float32)
    fa, ta = make_problem(pkg, init, target, res, size, cam, light, P, "l2", history=8)
    for _ in range(5):
        fa.step()
    fb, _ = make_problem(pkg, init, target, res, size, cam, light, P, "l2", history=8)
    for _ in range(2):
        fb.step()
    sd = fb.state_dict()
    fc, tc = make_problem(pkg, init, target, res, size, cam, light, P, "l2", history=8)
    fc.load_state_dict(sd)
    for _ in range(3):
        fc.step()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ta.cpu().numpy(), tc.cpu().numpy())
    la = fa.losses()
    hc = fc.losses()
    # the resumed run records steps 3..5 at their own indices
    assert len(hc) == 5
    np.testing.assert_array_equal(hc[2:5], la[2:5])
    assert np.all(hc[:2] == 0.0)
    assert fc.read_state()["step"] == 5.0


def test_gradient_against_fp64_finite_differences(pkg):
    """Coarse-to-fine check: 64 random pixels x 9 parameters, FP64 central
    differences (eps 1e-4) of the oracle loss; kinks are skipped."""
    res, V = 256, 9
    gt, _, size, cam, light, P = scene(pkg, res, V, seed=2)
    init = pkg.synthetic.init_textures(gt, seed=1)
    target = M.render(gt, res, size, cam, light, P, np.float32)
    fs, _ = make_problem(pkg, init, target, res, size, cam, light, P, "log_l2")
    fs.step()
    torch.cuda.synchronize()
    g_gpu = fs.exp_avg[0].cpu().numpy().astype(np.float64) / np.float64(np.float32(0.1))
    N = V * 3 * res * res
    rng = np.random.default_rng(2)
    h = 1e-4
    checked = 0
    for _ in range(64):
        i, j = rng.integers(0, res, 2)
        rows = slice(i, i + 1)
        tgt_row = target[:, :, rows]

        def L(t):
            out = M.render(np.clip(t, -1, 1), res, size, cam, light, P, np.float64, rows=rows)
            r = np.log(out) - np.log(np.clip(tgt_row.astype(np.float64), M.target_floor(np.float64), 1))
            return float((r * r).sum()) / N

        for c in range(9):
            if abs(abs(init[c, i, j]) - 1) < 2 * h:
                continue
            tp, tm = init.astype(np.float64), init.astype(np.float64)
            tp[c, i, j] += h
            tm[c, i, j] -= h
            L0, Lp, Lm = L(init.astype(np.float64)), L(tp), L(tm)
            fd = (Lp - Lm) / (2 * h)
            fwd, bwd = (Lp - L0) / h, (L0 - Lm) / h
            if abs(fwd - bwd) > 0.05 * max(abs(fwd), abs(bwd)) + 1e-30:
                continue   # a mask kink or strong curvature inside [t-h, t+h]
            scale = np.abs(g_gpu[c]).max()
            assert abs(g_gpu[c, i, j] - fd) <= 2e-3 * abs(fd) + 1e-6 * scale, (c, i, j, g_gpu[c, i, j], fd)
            checked += 1
    assert checked >= 400


def test_staged_256_512_1024(pkg):
    """Coarse to fine: 256^2 -> 512^2 -> 1024^2, 9 views, 20 iterations per stage,
    each stage starting from the bilinearly upsampled previous maps
    (svbrdf_upsample2x_bilinear, F.interpolate(mode='bilinear',
    align_corners=False) semantics); the material is
    one 1024^2 ground truth (one albedo / specular / roughness per channel,
    detail in the normals), box-averaged for the coarse stages.  The loss
    falls in every stage, every stage ends closer to its ground truth than
    it started, and every later stage starts below the 256^2 initial loss
    (the upsampled coarse solution is a better start than the noisy one)."""
    size = pkg.synthetic.PLANE_SIZE_CM
    cam, light = pkg.synthetic.flash_grid()
    P = pkg.synthetic.light_power()
    gt1024 = pkg.synthetic.ground_truth_textures(1024, seed=0, albedo="material")
    tex = None
    first = None
    for res in (256, 512, 1024):
        gt = pkg.synthetic.downsample(gt1024, 1024 // res)
        cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
        r = pkg.Microfacet(res, 9, size, cl, DEV)
        with torch.no_grad():
            target = r.eval(torch.tensor(gt, device=DEV)[None])
        if tex is None:
            tex = torch.tensor(pkg.synthetic.init_textures(gt, seed=1), device=DEV)[None].contiguous()
        else:
            tex = pkg.upsample_textures(tex)
        start_err = (tex[0] - torch.tensor(gt, device=DEV)).abs().mean().item()
        fs = pkg.FusedStep(r, tex, target, 0.01, loss="log_l2", history=20)
        for _ in range(20):
            fs.step()
        L = fs.losses()
        assert np.all(np.isfinite(L)) and L[-1] < L[0], (res, L)
        if first is None:
            first = L[0]
        else:
            assert L[0] < first, (res, L[0], first)
        end_err = (tex[0].clamp(-1, 1) - torch.tensor(gt, device=DEV)).abs().mean().item()
        assert end_err < start_err, (res, start_err, end_err)


# ---------------------------------------------------------------------------
# One iteration split in its two parts: the gradient the kernel computes (norm-
# wise against torch autograd in FP32 and FP64) and the Adam update it applies
# with that gradient (elementwise, every parameter, no outlier allowance).

def load_adam_state(pkg, fs, p, m, v, k):
    """Puts a FusedStep at the state torch.optim.Adam holds after k steps:
    maps p, moments m and v, step count k and the beta powers beta^(k + 1)
    of the next step (bias corrections as torch computes them, in double)."""
    dt = pkg.optim._STATE_DTYPE
    st = np.zeros(1, dt)
    b1p, b2p = fs.beta1 ** (k + 1), fs.beta2 ** (k + 1)
    st["beta1_pow"], st["beta2_pow"] = b1p, b2p
    st["inv_bias_correction1"] = 1.0 / (1.0 - b1p)
    st["inv_bias_correction2_sqrt"] = np.float32(1.0 / np.sqrt(1.0 - b2p))
    st["step"] = k
    raw = np.zeros(fs.state.numel() * 4, np.uint8)
    raw[:dt.itemsize] = st.view(np.uint8)
    shape = fs.textures.shape
    as_dev = lambda a: torch.as_tensor(np.asarray(a, np.float32)).reshape(shape).to(DEV)  # noqa: E731
    fs.load_state_dict({"textures": as_dev(p), "exp_avg": as_dev(m), "exp_avg_sq": as_dev(v),
                        "state": torch.from_numpy(raw.view(np.float32).copy()).to(DEV), "light_pow": fs.light_pow})


def kernel_gradient(pkg, fs, p):
    """The gradient the fused kernel computes at maps p: one step from zero
    moments at step 0, after which Adam's first moment is exactly
    fl(0.1 g); also the loss the kernel records for p."""
    z = np.zeros_like(p)
    load_adam_state(pkg, fs, p, z, z, 0)
    fs.step()
    torch.cuda.synchronize()
    g = fs.exp_avg.reshape(p.shape).cpu().numpy().astype(np.float64) / np.float64(np.float32(0.1))
    return g, fs.loss()


def check_adam_step(p0, m0, v0, g, k, p1, m1, v1, lr=0.01, b1=0.9, b2=0.999, eps=1e-8, label=""):
    """Every parameter of one step against Adam in float64 (torch's
    single-tensor order, torch/optim/adam.py:431-547) applied to the gradient
    g the kernel used: the moments within a few float32 roundings of their
    operands, the parameter within a few ulp of the update (the kernel takes
    the square root and the reciprocal from the hardware's approximate
    instructions, within 1 ulp each) plus the final rounding."""
    p0, m0, v0, g = (np.asarray(a, np.float64) for a in (p0, m0, v0, g))
    p1, m1, v1 = (np.asarray(a, np.float64) for a in (p1, m1, v1))
    w1 = np.float64(np.float32(1 - b1))
    m_exp = m0 + w1 * (g - m0)
    v_exp = v0 * np.float64(np.float32(b2)) + np.float64(np.float32(1 - b2)) * g * g
    u = 2.0 ** -24
    em = np.abs(m1 - m_exp) - 6 * u * (np.abs(m0) + np.abs(w1 * g)) - 1e-37
    ev = np.abs(v1 - v_exp) - 6 * u * (np.abs(v0) + 1e-3 * g * g) - 1e-37
    bc1, bc2 = 1 - b1 ** (k + 1), 1 - b2 ** (k + 1)
    step = np.float64(np.float32(-lr / bc1))
    upd = step * m1 / (np.sqrt(v1) * np.float64(np.float32(1 / np.sqrt(bc2))) + np.float64(np.float32(eps)))
    p_exp = p0 + upd
    ep = np.abs(p1 - p_exp) - 8 * u * np.abs(upd) - u * np.maximum(np.abs(p0), np.abs(p1)) - 1e-37
    for name, e in (("exp_avg", em), ("exp_avg_sq", ev), ("param", ep)):
        bad = e > 0
        assert not bad.any(), (f"{label} step {k + 1}: {int(bad.sum())} {name} values outside the bound, worst "
                               f"excess {e.max():.3e} at {np.unravel_index(e.argmax(), e.shape)}")


@pytest.mark.parametrize("loss,mode", [("log_l2", M.LOSS_LOG_L2), ("l2", M.LOSS_L2)])
@pytest.mark.parametrize("coloc", [True, False])
def test_first_step_full_maps_1024(pkg, loss, mode, coloc):
    """The headline shape, 1024^2 x 9 views: the fused kernel's first-step
    gradient over the FULL maps against torch autograd of the eager
    restatement of the reference renderer in FP32 and FP64 (run on the GPU),
    per channel under tests/parity.py's criteria; the loss within 1e-5 of
    FP64; and every parameter, moment and second moment of the update
    against Adam in float64 with that gradient.  Co-located flashes (the
    reference's capture geometry) and lights offset from their cameras (the
    general path)."""
    res, V = 1024, 9
    gt, _, size, cam, light, P = scene(pkg, res, V, seed=21, coloc=coloc)
    init = pkg.synthetic.init_textures(gt, seed=1)
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    with torch.no_grad():
        target = MT.MicrofacetEager(res, V, size, cl, DEV, torch.float32).eval(torch.tensor(gt, device=DEV)[None])
    fs, tex = make_problem(pkg, init, target.cpu().numpy(), res, size, cam, light, P, loss)
    fs.step()
    torch.cuda.synchronize()
    g_gpu = fs.exp_avg[0].cpu().numpy().astype(np.float64) / np.float64(np.float32(0.1))
    L32, g32 = eager_loss_grad(init, target, res, size, cam, light, P, mode, torch.float32, DEV)
    L64, g64, kinks = eager_loss_grad(init, target, res, size, cam, light, P, mode, torch.float64, DEV, kinks=True)
    torch.cuda.empty_cache()
    errs = parity.check_gradient(g_gpu, g32, g64, f"1024^2 {loss} coloc={coloc}", exclude=kinks)
    print(f"1024^2 x 9 {loss} coloc={coloc}: per-channel |g - g64|/|g64| max {max(errs):.2e}")
    assert abs(fs.loss() - L64) <= 1e-5 * abs(L64)
    z = np.zeros_like(init)
    check_adam_step(init, z, z, g_gpu, 0, tex[0].cpu().numpy(), fs.exp_avg[0].cpu().numpy(),
                    fs.exp_avg_sq[0].cpu().numpy(), label=f"1024^2 {loss}")


def reference_trajectory(res, V, size, cam, light, P, init, target, steps, mode, device):
    """The reference loop (src/svbrdf.py:56-71; eager restatement + torch
    Adam, single-tensor) with its state before every step: maps, moments,
    its own FP32 gradient and loss."""
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    r = MT.MicrofacetEager(res, V, size, cl, device, torch.float32)
    t = torch.tensor(init, device=device)[None].clone().requires_grad_(True)
    opt = torch.optim.Adam([t], lr=0.01, betas=(0.9, 0.999), foreach=False)
    tg = torch.as_tensor(target).to(device)
    states = []
    for k in range(steps):
        st = opt.state.get(t)
        p = t.detach()[0].cpu().numpy().copy()
        m = st["exp_avg"][0].cpu().numpy().copy() if st else np.zeros_like(p)
        v = st["exp_avg_sq"][0].cpu().numpy().copy() if st else np.zeros_like(p)
        L = MT.image_loss(r.eval(t.clamp(-1, 1)), tg, mode)
        opt.zero_grad()
        L.backward()
        states.append((p, m, v, t.grad[0].cpu().numpy().copy(), L.item()))
        opt.step()
    return states, t.detach()[0].cpu().numpy()


def check_teacher_forced(pkg, fs, states, res, size, cam, light, P, target, mode, device, label):
    """At every state of the reference trajectory: the kernel's gradient and
    loss at the reference's maps (norm-wise per channel against the
    reference's FP32 gradient and the FP64 gradient there), then one kernel
    step from the reference's full Adam state, checked elementwise."""
    worst = 0.0
    for k, (p, m, v, g32, _L32) in enumerate(states):
        g_gpu, L_gpu = kernel_gradient(pkg, fs, p)
        L64, g64, kinks = eager_loss_grad(p, target, res, size, cam, light, P, mode, torch.float64, device, kinks=True)
        worst = max(worst, max(parity.check_gradient(g_gpu, g32, g64, f"{label} step {k + 1}", exclude=kinks)))
        assert abs(L_gpu - L64) <= 1e-5 * abs(L64), (k, L_gpu, L64)
        load_adam_state(pkg, fs, p, m, v, k)
        fs.step()
        torch.cuda.synchronize()
        check_adam_step(p, m, v, g_gpu, k, fs.textures.reshape(p.shape).cpu().numpy(),
                        fs.exp_avg.reshape(p.shape).cpu().numpy(), fs.exp_avg_sq.reshape(p.shape).cpu().numpy(),
                        label=label)
    return worst


@pytest.mark.parametrize("loss", ["l2", "logl2"])
def test_golden_trajectory_step_by_step(pkg, golden_dir, loss):
    """The golden 5-step runs of the reference renderer + torch.optim.Adam
    (R = 8): the eager restatement reproduces their maps (bit for bit on the
    host that wrote them), and at each of its states the kernel's gradient and its Adam step
    are checked as one iteration's two parts -- no parameter may differ by
    more than the update's rounding."""
    z = np.load(golden_dir / f"optim_{loss}_r8.npz")
    res, size, lr = int(z["res"]), float(z["size"]), float(z["lr"])
    mode = M.LOSS_L2 if loss == "l2" else M.LOSS_LOG_L2
    assert lr == 0.01
    states, final = reference_trajectory(res, z["cam"].shape[0], size, z["cam"], z["light"], z["light_pow0"],
                                         z["init"], z["target"], 5, mode, "cpu")
    # bit for bit where the golden file was written (tests/test_oracle.py pins
    # that); on a host whose CPU kernels take another vector path a few values
    # differ in the last bits, which Adam carries on over the steps
    for k, got in enumerate([st[0] for st in states[1:]] + [final]):
        ref = z["tex_textures"][k]
        d = np.abs(got - ref)
        assert d.max() <= 2e-6 and d.mean() <= 5e-8, (k, float(d.max()), float(d.mean()), np.mean(d == 0))
    fs, _ = make_problem(pkg, z["init"], z["target"], res, size, z["cam"], z["light"], z["light_pow0"], loss)
    check_teacher_forced(pkg, fs, states, res, size, z["cam"], z["light"], z["light_pow0"], z["target"], mode,
                         "cpu", f"golden {loss}")


def test_config0_trajectory_step_by_step(pkg):
    """Config 0 (256^2, 9 views, log-L2, 20 Adam iterations at lr 0.01): at
    each of the 20 states of the reference loop the kernel's gradient over
    the full maps (FP32 / FP64 autograd, on the GPU) and every value of its
    Adam step."""
    res = 256
    gt, _, size, cam, light, P = scene(pkg, res, 9, seed=0)
    init = pkg.synthetic.init_textures(gt, seed=1)
    target = M.render(gt, res, size, cam, light, P, np.float32)
    states, _ = reference_trajectory(res, 9, size, cam, light, P, init, target, 20, M.LOSS_LOG_L2, DEV)
    fs, _ = make_problem(pkg, init, target, res, size, cam, light, P, "log_l2")
    worst = check_teacher_forced(pkg, fs, states, res, size, cam, light, P, target, M.LOSS_LOG_L2, DEV, "config0")
    print(f"config 0: worst per-channel |g - g64|/|g64| over 20 states {worst:.2e}")


@pytest.mark.parametrize("ch", [0, 3, 5, 6])
def test_nan_texel_in_fused_step(pkg, ch):
    """A NaN texel in the maps (diffuse, normal x, roughness or specular).
    As in the reference: the loss is NaN, the texel's own gradient is 0 (the
    clamp on read masks it), so the texel stays NaN and its moments 0.  The
    rest of that pixel's gradient: the kernel masks the NaN rendered value's
    residual (v_med3_f32 keeps the clamped value inside [eps, 1]) and
    multiplies the masked zero by finite factors where it can, so its NaN
    entries are a subset of the reference's (the reference's autograd
    multiplies the masked zero by the NaN factors of the product rule; a NaN
    diffuse texel makes the reference's normal gradient NaN, the kernel's
    stays finite), and the entries finite in both agree.  Every other pixel
    is bit-identical to a run without the NaN."""
    res, V = 64, 9
    gt, _, size, cam, light, P = scene(pkg, res, V, seed=3)
    init = pkg.synthetic.init_textures(gt, seed=1)
    target = M.render(gt, res, size, cam, light, P, np.float32)
    i, j = 17, 40
    bad = init.copy()
    bad[ch, i, j] = np.nan
    fa, ta = make_problem(pkg, bad, target, res, size, cam, light, P, "log_l2")
    fb, tb = make_problem(pkg, init, target, res, size, cam, light, P, "log_l2")
    fa.step()
    fb.step()
    torch.cuda.synchronize()
    assert np.isnan(fa.loss()) and np.isfinite(fb.loss())
    pa, pb = ta[0].cpu().numpy(), tb[0].cpu().numpy()
    ma, mb = fa.exp_avg[0].cpu().numpy(), fb.exp_avg[0].cpu().numpy()
    va, vb = fa.exp_avg_sq[0].cpu().numpy(), fb.exp_avg_sq[0].cpu().numpy()
    other = np.ones((res, res), bool)
    other[i, j] = False
    for a, b in ((pa, pb), (ma, mb), (va, vb)):
        assert np.array_equal(a[:, other], b[:, other])
    assert np.isnan(pa[ch, i, j]) and ma[ch, i, j] == 0 and va[ch, i, j] == 0
    # the reference's gradient at that pixel (FP64 autograd of the eager restatement)
    _, g64 = eager_loss_grad(bad, target, res, size, cam, light, P, M.LOSS_LOG_L2, torch.float64)
    g_gpu = ma[:, i, j].astype(np.float64) / np.float64(np.float32(0.1))
    ref = g64[:, i, j]
    assert ref[ch] == 0.0 and g_gpu[ch] == 0.0
    assert np.all(np.isnan(ref[np.isnan(g_gpu)])), (g_gpu, ref)
    both = np.isfinite(ref) & np.isfinite(g_gpu)
    np.testing.assert_allclose(g_gpu[both], ref[both], rtol=1e-4, atol=1e-12)
    # the parameters follow their gradients: NaN only where the gradient is
    assert np.array_equal(np.isnan(pa[:, i, j]), np.isnan(g_gpu) | (np.arange(9) == ch))
