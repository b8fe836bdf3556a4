"""Batched persistent step: many materials of uneven size and view count in
one launch.  Each material's per-pixel arithmetic is the single-material
kernel's (same source, separate instantiation: instruction selection may
contract a multiply-add differently, measured at ~0.2 % of values one ulp
apart after 3 steps), so its parameters must match running it alone through
FusedStep bit for bit on >= 99 % of the values, to 1e-5 relative on the
parameters and 1e-4 relative on the second moments (whose one-ulp gradient
differences are squared), whatever the launch shape; losses are reduced per
tile instead of per block, so they agree to rounding.

Against the reference: the config-3 batch (64 materials, 24.6 Mpx) and a
batch of uneven shapes, in several launch forms, each material's first step
checked over its full maps as one iteration of the reference loop on that
material alone (gradient against FP32 / FP64 torch autograd of the eager
restatement, loss, every value of the Adam update)."""
import numpy as np
import pytest
import torch

from oracle import microfacet_np as M
from oracle import microfacet_torch as MT
from tests import parity
from tests.test_gpu_step import check_adam_step, eager_loss_grad

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")

# (res, views, co-located)
SHAPES = [(64, 9, True), (33, 4, True), (100, 16, True), (16, 1, True), (48, 5, False), (128, 12, True)]


def make_problems(pkg, shapes, seed=0):
    out = []
    for i, (res, V, coloc) in enumerate(shapes):
        gt = pkg.synthetic.ground_truth_textures(res, seed=seed + i)
        init = pkg.synthetic.init_textures(gt, seed=seed + 50 + i)
        rng = np.random.default_rng(seed + 200 + i)
        cam, light = pkg.synthetic.flash_grid(jitter=3.0, rng=rng, n_views=V)
        if not coloc:
            light = (light + rng.uniform(-2, 2, light.shape)).astype(np.float32)
        P = pkg.synthetic.light_power()
        size = pkg.synthetic.PLANE_SIZE_CM
        target = M.render(gt, res, size, cam, light, P, np.float32)
        out.append((res, V, size, cam, light, P, init, target))
    return out


def build(pkg, specs, kind, **kw):
    problems, texs = [], []
    for res, V, size, cam, light, P, init, target in specs:
        cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
        r = pkg.Microfacet(res, V, size, cl, DEV)
        tex = torch.tensor(init, device=DEV)[None].contiguous()
        problems.append((r, tex, torch.tensor(target, device=DEV)))
        texs.append(tex)
    if kind == "batch":
        return pkg.BatchedStep(problems, 0.01, **kw), texs
    return [pkg.FusedStep(r, t, y, 0.01, history=8, **kw) for r, t, y in problems], texs


@pytest.mark.parametrize("cfg", [
    {"tile_pixels": 256, "pixels_per_thread": 1, "views_per_trip": 1, "blocks_per_cu": 1},
    {"tile_pixels": 4096, "pixels_per_thread": 1, "views_per_trip": 3, "blocks_per_cu": 0},
    {"tile_pixels": 1024, "pixels_per_thread": 2, "views_per_trip": 1, "blocks_per_cu": 2},
    {"tile_pixels": 16384, "pixels_per_thread": 2, "views_per_trip": 3, "blocks_per_cu": 0},
    {"tile_pixels": 256, "pixels_per_thread": 1, "views_per_trip": 1, "blocks_per_cu": -1},
    {"tile_pixels": 256, "pixels_per_thread": 1, "views_per_trip": 3, "blocks_per_cu": -1},
])
@pytest.mark.parametrize("loss", ["log_l2", "l2"])
def test_batch_matches_single_material_steps(pkg, cfg, loss):
    specs = make_problems(pkg, SHAPES)
    batch, btex = build(pkg, specs, "batch", loss=loss, history=8, config=cfg)
    singles, stex = build(pkg, specs, "single", loss=loss)
    for _ in range(3):
        batch.step()
        for f in singles:
            f.step()
    torch.cuda.synchronize()
    for i, (b, s) in enumerate(zip(btex, stex)):
        pairs = ((b, s, 1e-5, 1e-6), (batch.exp_avg_sq[i], singles[i].exp_avg_sq, 1e-4, 1e-14))
        for x, y, rtol, atol in pairs:
            x, y = x.cpu().numpy(), y.cpu().numpy()
            frac = np.mean(x == y)
            far = ~np.isclose(x, y, rtol=rtol, atol=atol)
            info = (i, SHAPES[i], frac, int(far.sum()), float(np.abs(x - y).max()))
            assert frac >= 0.99, info
            # a gradient within an ulp of zero may flip sign between the two
            # instantiations; Adam then moves that one value by up to
            # lr*(1-b1)/sqrt(1-b2) per step the other way, nothing else
            assert far.sum() <= max(2, 1e-4 * x.size), info
            assert np.abs(x - y).max() <= 3 * 2 * 0.01 * 0.1 / np.sqrt(1e-3) + 1e-6, info
    hist = batch.loss_history()
    assert hist.shape == (3, len(SHAPES))
    for i, f in enumerate(singles):
        np.testing.assert_allclose(hist[:, i], f.losses(), rtol=1e-5)
    assert batch.read_state()["step"] == 3.0


def test_batch_graph_replay_and_determinism(pkg):
    specs = make_problems(pkg, SHAPES, seed=7)
    a, ta = build(pkg, specs, "batch", history=4)
    b, tb = build(pkg, specs, "batch", history=4)
    for _ in range(3):
        a.step()
    b.capture()
    for _ in range(3):
        b.replay()
        _ = (tb[0].sum() * 0).item()
    torch.cuda.synchronize()
    for x, y in zip(ta, tb):
        np.testing.assert_array_equal(x.cpu().numpy(), y.cpu().numpy())
    np.testing.assert_array_equal(a.loss_history(), b.loss_history())


def test_batch_own_textures_and_multi_step_graphs(pkg):
    """Maps at the padded plane stride (own_textures) and graphs of several
    iterations: the same arithmetic as one eager launch per step on the
    caller's tensors."""
    specs = make_problems(pkg, SHAPES, seed=9)
    a, ta = build(pkg, specs, "batch", history=8)
    b, tb = build(pkg, specs, "batch", history=8, own_textures=True)
    for _ in range(6):
        a.step()
    b.capture(3)
    b.replay()
    b.replay()
    torch.cuda.synchronize()
    for t, spec in zip(tb, specs):
        assert np.array_equal(t[0].cpu().numpy(), spec[6])      # untouched until write_back
    b.write_back()
    torch.cuda.synchronize()
    for i, (x, y) in enumerate(zip(ta, tb)):
        np.testing.assert_array_equal(x.cpu().numpy(), y.cpu().numpy())
        np.testing.assert_array_equal(a.exp_avg_sq[i].cpu().numpy(), b.exp_avg_sq[i].cpu().numpy())
    np.testing.assert_array_equal(a.loss_history(), b.loss_history())
    assert b.steps_launched == 6 and b.read_state()["step"] == 6.0


def test_autotune_restores_state_and_picks_a_candidate(pkg, tmp_path, monkeypatch):
    monkeypatch.setenv("SVBRDF_AMD_CACHE", str(tmp_path))
    specs = make_problems(pkg, SHAPES[:4], seed=3)
    b, tex = build(pkg, specs, "batch", history=4)
    b.step()
    torch.cuda.synchronize()
    before = [t.cpu().numpy().copy() for t in tex]
    st_before = b.read_state().copy()
    cands = pkg.BatchedStep.candidates()[:6]
    cfg, timings = b.autotune(candidates=cands, iters=2)
    assert len(timings) >= 1 and all(t > 0 for t in timings.values())
    assert cfg in [dict(c, blocks_per_cu=cfg["blocks_per_cu"]) for c in cands]
    after = [t.cpu().numpy() for t in tex]
    for x, y in zip(before, after):
        np.testing.assert_array_equal(x, y)
    assert b.read_state()["step"] == st_before["step"]
    # the cache file now holds the winner and a second autotune reads it back
    cfg2, t2 = b.autotune()
    assert cfg2 == cfg and t2 == {}
    b.step()
    torch.cuda.synchronize()
    assert b.read_state()["step"] == 2.0


# ---------------------------------------------------------------------------
# Against the reference: each material's first step in the batch is one
# iteration of the reference loop (src/svbrdf.py:56-71) on that material.

PERSISTENT = {"tile_pixels": 4096, "pixels_per_thread": 2, "views_per_trip": 3, "blocks_per_cu": 0}
GRID = {"tile_pixels": 256, "pixels_per_thread": 1, "views_per_trip": 1, "blocks_per_cu": -1}


def first_step_in_batch(pkg, problems, cfg, loss):
    """One batched step; per material (maps, exp_avg, exp_avg_sq) after it
    (host copies) and the losses the batch recorded for it."""
    bs = pkg.BatchedStep([(r, t.clone(), y) for r, t, y in problems], 0.01, loss=loss, config=cfg, history=1)
    assert (bs.grid == 0) == (cfg["blocks_per_cu"] < 0)
    bs.step()
    torch.cuda.synchronize()
    out = [(t.reshape(9, r.res, r.res).cpu().numpy(), m.reshape(9, r.res, r.res).cpu().numpy(),
            v.reshape(9, r.res, r.res).cpu().numpy()) for r, t, m, v in zip(bs.renderers, bs.textures, bs.exp_avg,
                                                                            bs.exp_avg_sq)]
    losses = bs.losses()
    del bs
    torch.cuda.empty_cache()
    return out, losses


def check_batch_against_eager(pkg, specs, problems, results, loss, mode, label):
    """Per material: the kernel's gradient (from Adam's first moment) against
    FP32 / FP64 torch autograd of the eager restatement (on the GPU), under
    tests/parity.py's criteria; its loss within 1e-5 of FP64; every value of
    its Adam step against float64 Adam with that gradient."""
    worst = 0.0
    for i, spec in enumerate(specs):
        res, V, size, cam, light, P, init, target = spec
        L32, g32 = eager_loss_grad(init, target, res, size, cam, light, P, mode, torch.float32, DEV)
        L64, g64, kinks = eager_loss_grad(init, target, res, size, cam, light, P, mode, torch.float64, DEV,
                                          kinks=True)
        for form, (per_mat, losses) in results.items():
            p1, m1, v1 = per_mat[i]
            g = m1.astype(np.float64) / np.float64(np.float32(0.1))
            tag = f"{label} {form} material {i} ({res}^2 x {V})"
            worst = max(worst, max(parity.check_gradient(g, g32, g64, tag, exclude=kinks)))
            assert abs(losses[i] - L64) <= 1e-5 * abs(L64), (tag, losses[i], L64)
            z = np.zeros_like(init)
            check_adam_step(init, z, z, g, 0, p1, m1, v1, label=tag)
        torch.cuda.empty_cache()
    return worst


@pytest.mark.timeout(900)
def test_config3_batch_against_the_reference(pkg):
    """Config 3 as bench.py runs it: 64 materials (24 at 256^2, 24 at 512^2,
    16 at 1024^2), views drawn from 4..16, flash positions jittered +-3 cm
    around the 3x3 grid (bench.batch_materials, seed 3), log-L2.  One batched
    step in each launch form -- one block per 256-pixel tile (the autotuned
    form on this batch) and a persistent grid drawing 4,096-pixel tiles, two
    pixels per thread, three views per loop trip -- and every material's
    update checked over its full maps against one iteration of the
    reference loop on that material alone."""
    import bench
    syn = pkg.synthetic
    mats = bench.batch_materials(pkg, DEV)
    assert sorted(m[0] for m in mats) == [256] * 24 + [512] * 24 + [1024] * 16
    assert min(m[1] for m in mats) >= 4 and max(m[1] for m in mats) <= 16
    P, size = syn.light_power(), syn.PLANE_SIZE_CM
    specs, problems = [], []
    for res, V, gt, init, cam, light in mats:
        cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
        with torch.no_grad():
            tgt = MT.MicrofacetEager(res, V, size, cl, DEV, torch.float32).eval(torch.tensor(gt, device=DEV)[None])
        r = pkg.Microfacet(res, V, size, cl, DEV)
        problems.append((r, torch.tensor(init, device=DEV)[None].contiguous(), tgt))
        specs.append((res, V, size, cam, light, P, init, tgt))
    results = {name: first_step_in_batch(pkg, problems, cfg, "log_l2")
               for name, cfg in (("grid", GRID), ("persistent", PERSISTENT))}
    worst = check_batch_against_eager(pkg, specs, problems, results, "log_l2", MT.LOSS_LOG_L2, "config 3")
    print(f"config 3: worst per-channel |g - g64|/|g64| over 64 materials x 2 forms {worst:.2e}")


@pytest.mark.parametrize("loss,mode", [("log_l2", MT.LOSS_LOG_L2), ("l2", MT.LOSS_L2)])
def test_uneven_batch_against_the_reference(pkg, loss, mode):
    """Odd sizes (33^2, 100^2: tiles that end inside a tile), one view,
    sixteen views and a material whose lights sit off its cameras (the
    general path), in three launch forms, against the reference loop on
    each material alone."""
    specs = make_problems(pkg, SHAPES, seed=30)
    problems = []
    for res, V, size, cam, light, P, init, target in specs:
        cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
        problems.append((pkg.Microfacet(res, V, size, cl, DEV), torch.tensor(init, device=DEV)[None].contiguous(),
                         torch.tensor(target, device=DEV)))
    forms = (("grid", GRID), ("persistent", PERSISTENT),
             ("persistent-1px", {"tile_pixels": 1024, "pixels_per_thread": 1, "views_per_trip": 1,
                                 "blocks_per_cu": 1}))
    results = {name: first_step_in_batch(pkg, problems, cfg, loss) for name, cfg in forms}
    check_batch_against_eager(pkg, specs, problems, results, loss, mode, "uneven batch")


def eager_light_grad(tex, target, res, size, cam, light, P, mode, dtype):
    """dL/dlight_pow of the reference loop's loss through the eager
    restatement (torch autograd on the CPU)."""
    p = torch.tensor(P, dtype=dtype, requires_grad=True)
    r = MT.MicrofacetEager(res, cam.shape[0], size, [torch.tensor(cam), torch.tensor(light), p], "cpu", dtype)
    t = torch.as_tensor(tex).to(dtype)
    L = MT.image_loss(r.eval(t[None].clamp(-1, 1)), torch.as_tensor(target).to(dtype), mode)
    L.backward()
    return p.grad.numpy().astype(np.float64)


@pytest.mark.parametrize("cfg", [
    {"tile_pixels": 256, "pixels_per_thread": 1, "views_per_trip": 1, "blocks_per_cu": -1},
    {"tile_pixels": 4096, "pixels_per_thread": 1, "views_per_trip": 3, "blocks_per_cu": 0},
    {"tile_pixels": 1024, "pixels_per_thread": 2, "views_per_trip": 1, "blocks_per_cu": 2},
])
@pytest.mark.parametrize("loss", ["log_l2", "l2"])
def test_batch_light_matches_single_material_steps(pkg, cfg, loss):
    """optim_light=True in the batch: every material's light power takes its
    own Adam step (src/svbrdf.py:48-50,58-59 per material).  Three steps
    against FusedStep(optim_light=True) run on each material alone (the
    path the reference's golden light trajectories pin): light powers to 1e-5
    relative (the batch sums dL/dP per tile, the single kernel per block, so
    the two double sums may round one ulp apart), maps and second moments as
    in test_batch_matches_single_material_steps, losses to rounding."""
    specs = make_problems(pkg, SHAPES, seed=3)
    batch, btex = build(pkg, specs, "batch", loss=loss, history=8, config=cfg, optim_light=True)
    singles, stex = build(pkg, specs, "single", loss=loss, optim_light=True)
    for _ in range(3):
        batch.step()
        for f in singles:
            f.step()
    torch.cuda.synchronize()
    for i, f in enumerate(singles):
        np.testing.assert_allclose(batch.light_pow[i].cpu().numpy(), f.light_pow.cpu().numpy(), rtol=1e-5,
                                   err_msg=str((i, SHAPES[i])))
        assert not np.array_equal(batch.light_pow[i].cpu().numpy(), specs[i][5]), "light power did not move"
    for i, (b, s) in enumerate(zip(btex, stex)):
        for x, y, rtol, atol in ((b, s, 1e-5, 1e-6), (batch.exp_avg_sq[i], singles[i].exp_avg_sq, 1e-4, 1e-14)):
            x, y = x.cpu().numpy(), y.cpu().numpy()
            far = ~np.isclose(x, y, rtol=rtol, atol=atol)
            info = (i, SHAPES[i], float(np.mean(x == y)), int(far.sum()), float(np.abs(x - y).max()))
            assert np.mean(x == y) >= 0.99, info
            assert far.sum() <= max(2, 1e-4 * x.size), info
            assert np.abs(x - y).max() <= 3 * 2 * 0.01 * 0.1 / np.sqrt(1e-3) + 1e-6, info
    hist = batch.loss_history()
    for i, f in enumerate(singles):
        np.testing.assert_allclose(hist[:, i], f.losses(), rtol=1e-5)


@pytest.mark.parametrize("loss,mode", [("l2", M.LOSS_L2), ("log_l2", M.LOSS_LOG_L2)])
def test_batch_light_first_step_against_the_reference(pkg, loss, mode):
    """One batched step with optim_light=True: each material's dL/dlight_pow,
    recovered exactly from its light Adam moment (m1 = 0.1 g), against torch
    autograd of the eager restatement in FP32 and FP64 (within 1e-4 of FP64,
    or no further than twice the reference FP32's own error), and the light
    power after the step against Adam's first step from that gradient
    (p1 = p0 - lr g / (|g| + eps))."""
    specs = make_problems(pkg, SHAPES, seed=5)
    batch, _ = build(pkg, specs, "batch", loss=loss, optim_light=True)
    batch.step()
    torch.cuda.synchronize()
    adam = batch.light_adam.cpu().numpy().astype(np.float64)
    for i, (res, V, size, cam, light, P, init, target) in enumerate(specs):
        g = adam[i, :3] / np.float64(np.float32(0.1))
        g64 = eager_light_grad(init, target, res, size, cam, light, P, mode, torch.float64)
        g32 = eager_light_grad(init, target, res, size, cam, light, P, mode, torch.float32)
        bound = np.maximum(1e-4 * np.abs(g64), 2 * np.abs(g32 - g64)) + 1e-30
        assert np.all(np.abs(g - g64) <= bound), (i, SHAPES[i], g, g64, g32)
        p1 = np.asarray(P, np.float64) - 0.01 * g / (np.abs(g) + 1e-8)
        np.testing.assert_allclose(batch.light_pow[i].cpu().numpy(), p1, rtol=1e-6, err_msg=str(i))
