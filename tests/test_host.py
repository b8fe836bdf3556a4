"""Host-side pieces that need no GPU: calibration loading, the synthetic
workloads, the batch planner's tile tables and bench's torch-autograd
baseline (which must compute the reference's images)."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import microfacet_np as M
from oracle import microfacet_torch as MT
from tests.conftest import ROOT


def test_load_calibration_selects_views_in_idx_order(pkg):
    import json
    from svbrdf_amd.svbrdf_io import calibration_arrays
    data = json.loads((ROOT / "tests" / "fixtures" / "calib_grid.json").read_text())
    assert data["im_size"] == 12.5 and data["idx"] == [3, 0, 2]
    cam, light, pow_ = calibration_arrays(data["camera_pos"], data["light_pos"], data["light_pow"], data["idx"])
    np.testing.assert_array_equal(cam, [[2, -2, 20], [-2, 2, 20], [-2, -2, 20]])
    np.testing.assert_array_equal(light[:, 2], [21, 21, 21])
    assert cam.dtype == light.dtype == pow_.dtype == np.float32 and pow_.tolist() == [900, 1000, 1100]
    assert cam.flags["C_CONTIGUOUS"] and light.flags["C_CONTIGUOUS"]
    with pytest.raises(ValueError):
        calibration_arrays(data["camera_pos"], data["light_pos"], [1.0, 2.0], data["idx"])


def test_synthetic_maps_follow_the_recipe(pkg):
    syn = pkg.synthetic
    t = syn.ground_truth_textures(128, seed=0)
    assert t.shape == (9, 128, 128) and t.dtype == np.float32
    u = (t + 1) / 2
    assert 0.05 <= u[0:3].min() and u[0:3].max() <= 0.9
    assert 0.02 <= u[6:9].min() and u[6:9].max() <= 0.3
    assert 0.1 <= u[5].min() and u[5].max() <= 0.9
    assert (t[3] ** 2 + t[4] ** 2).max() < 1          # unit normals: |(nx, ny)| < 1
    cam, light = syn.flash_grid()
    assert cam.shape == (9, 3) and np.array_equal(cam, light)
    assert cam[0].tolist() == [-8, 8, 30] and cam[8].tolist() == [8, -8, 30]
    init = syn.init_textures(t, seed=1)
    assert np.abs(init).max() <= 1 and 0.05 < np.abs(init - t).mean() < 0.1
    np.testing.assert_array_equal(syn.ground_truth_textures(64, seed=5), syn.ground_truth_textures(64, seed=5))


def test_config3_batch_shape():
    import bench
    sizes = [256] * 24 + [512] * 24 + [1024] * 16
    assert sum(r * r for r in sizes) == 24_641_536
    rng = np.random.default_rng(3)
    views = rng.integers(4, 17, len(sizes))
    assert views.min() >= 4 and views.max() <= 16
    assert bench.algorithmic_bytes_step(1024, 9) == 339_738_624


def test_batch_plan_orders_tiles_by_cost(pkg):
    from svbrdf_amd.batch import Material
    lib = pkg._lib.load()
    shapes = [(256, 5), (1024, 16), (13, 4), (512, 9)]
    mats = (Material * len(shapes))()
    for i, (r, v) in enumerate(shapes):
        mats[i] = Material(0x1000, 0x2000, 0x3000, 0x4000, 0x5000, 0x6000, 0x7000, r, v, 17.0, 0, 0, 0)
    n = lib.svbrdf_batch_tile_count(mats, len(shapes), 4096)
    assert n == 16 + 256 + 1 + 64
    order = (ctypes.c_uint * n)()
    tile_mat = (ctypes.c_uint * n)()
    assert lib.svbrdf_batch_plan(mats, len(shapes), 4096, order, tile_mat, n) == 0
    assert [(m.tile_begin, m.n_tiles) for m in mats] == [(0, 16), (16, 256), (272, 1), (273, 64)]
    order = list(order)
    assert sorted(order) == list(range(n))

    def cost(t):
        r, v = shapes[tile_mat[t]]
        first = (t - mats[tile_mat[t]].tile_begin) * 4096
        return min(4096, r * r - first) * v

    costs = [cost(t) for t in order]
    assert costs == sorted(costs, reverse=True)
    assert lib.svbrdf_batch_plan(mats, len(shapes), 4096, (ctypes.c_uint * 3)(), tile_mat, 3) == -3
    assert lib.svbrdf_batch_plan(mats, len(shapes), 1000, (ctypes.c_uint * n)(), tile_mat, n) == -1
    # plane strides below res^2 are refused, padded ones accepted
    mats[0].plane_stride = 256 * 256 - 1
    assert lib.svbrdf_batch_plan(mats, len(shapes), 4096, order_arr(n), tile_mat, n) == -1
    mats[0].plane_stride = lib.svbrdf_plane_stride(256)
    mats[0].texture_stride = 100
    assert lib.svbrdf_batch_plan(mats, len(shapes), 4096, order_arr(n), tile_mat, n) == -1
    mats[0].texture_stride = lib.svbrdf_plane_stride(256)
    assert lib.svbrdf_batch_plan(mats, len(shapes), 4096, order_arr(n), tile_mat, n) == 0


def order_arr(n):
    return (ctypes.c_uint * n)()


def test_batch_step_one_block_per_tile_needs_256_pixel_tiles(pkg):
    """grid_blocks = 0 (one block per tile) is refused on the host for other
    tile sizes or two pixels per thread, before anything is launched."""
    lib = pkg._lib.load()
    fake = ctypes.c_void_p(0x10000)
    ws = lib.svbrdf_batch_workspace_bytes(64)

    def step(tile, npix_t, grid):
        return lib.svbrdf_batch_step(fake, 1, fake, fake, 64, tile, npix_t, 1, grid, 1, 0, ctypes.c_double(0.01),
                                     ctypes.c_double(0.9), ctypes.c_double(0.999), ctypes.c_double(1e-8),
                                     fake, fake, None, 0, fake, ws, None)
    assert step(1024, 1, 0) == -1
    assert step(512, 2, 0) == -1
    assert step(256, 1, -1) == -1
    assert b"one block per tile" in lib.svbrdf_last_error() or b"grid_blocks" in lib.svbrdf_last_error()


def test_bench_eager_baseline_renders_the_reference_images():
    import bench
    rng = np.random.default_rng(0)
    res, V = 16, 4
    tex = rng.uniform(-0.9, 0.9, (1, 9, res, res)).astype(np.float32)
    tex[:, 3:5] *= 0.4
    cam = np.array([[x, y, 25.0] for x in (-5.0, 5.0) for y in (-5.0, 5.0)], np.float32)
    P = np.array([1000, 1200, 1400], np.float32)
    path = bench.EagerReferencePath(res, V, 17.0, cam, cam, P, "cpu")
    out = path.eval(torch.tensor(tex)).numpy()
    ref = M.render(tex[0], res, 17.0, cam, cam, P, np.float64)
    np.testing.assert_allclose(out, ref, rtol=2e-4, atol=1e-6)
    cl = [torch.tensor(cam), torch.tensor(cam), torch.tensor(P)]
    ref32 = MT.MicrofacetEager(res, V, 17.0, cl, "cpu").eval(torch.tensor(tex)).numpy()
    np.testing.assert_allclose(out, ref32, rtol=2e-4, atol=1e-6)


@pytest.mark.parametrize("epochs,save_every,sync_every", [(1, 100, 0), (2, 100, 0), (100, 100, 0),
                                                          (250, 100, 0), (1000, 100, 64), (37, 5, 7)])
def test_epoch_segments_cover_the_loop_and_end_at_save_points(pkg, epochs, save_every, sync_every):
    segs = pkg.optim.epoch_segments(epochs, save_every, sync_every)
    # contiguous, in order, covering every epoch once
    assert segs[0][0] == 0 and segs[-1][1] == epochs - 1
    assert all(a[1] + 1 == b[0] for a, b in zip(segs, segs[1:]))
    saves = [last + 1 for _, last, save in segs if save]
    # the reference saves after epoch 1, every 100th and the last (src/svbrdf.py:74)
    want = sorted({1, epochs} | set(range(save_every, epochs + 1, save_every)))
    assert saves == want
    if sync_every:
        ends = {last + 1 for _, last, _ in segs}
        assert set(range(sync_every, epochs + 1, sync_every)) <= ends
        assert all(last - first + 1 <= sync_every for first, last, _ in segs)
    # without saving or syncing the whole run is one segment
    assert pkg.optim.epoch_segments(epochs, save_every, 0, saving=False) == [(0, epochs - 1, False)]


def test_save_loss_writes_the_reference_plot(pkg, tmp_path):
    """src/optimization.py:35-42: an 8 x 4 inch log(1 + loss) plot; drawn
    without touching pyplot's global state."""
    from PIL import Image
    import matplotlib
    backend = matplotlib.get_backend()
    pkg.SvbrdfOptim.save_loss([[0.5, 0.2, 0.1], [0.05, 0.04, 0.03]], ["image loss", "feature loss"],
                              tmp_path / "loss.jpg", 1000)
    with Image.open(tmp_path / "loss.jpg") as im:
        assert im.size == (800, 400)
    assert matplotlib.get_backend() == backend


def _optim_on_host(pkg, res):
    """SvbrdfOptim's initialisers are torch ops on self.device; on the host
    they run without the library (the optimiser itself needs the GPU)."""
    from types import SimpleNamespace
    return pkg.SvbrdfOptim("cpu", SimpleNamespace(res=res))


def test_init_from_const_encodings(pkg):
    """src/svbrdf.py:23-30: u = dif / spe / rgh stored as t = 2u - 1 in the
    channel order of src/svbrdf.py:166 (diffuse 0:3, normal xy 3:5, roughness
    5, specular 6:9), normal xy = 0 (a flat normal); a leaf with grad."""
    o = _optim_on_host(pkg, 8)
    o.init_from_const()
    t = o.textures
    assert t.shape == (1, 9, 8, 8) and t.dtype == torch.float32 and t.requires_grad and t.is_leaf
    x = t.detach()[0].numpy()
    for c in range(9):
        assert np.unique(x[c]).size == 1, c
    v = x[:, 0, 0].astype(np.float64)
    u = (v + 1) / 2
    np.testing.assert_allclose(u[0:3], 0.5, atol=1e-7)
    np.testing.assert_allclose(u[6:9], 0.04, atol=1e-7)
    np.testing.assert_allclose(u[5], 0.2, atol=1e-7)
    assert v[3] == 0.0 and v[4] == 0.0
    # 0.5 -> exactly 0; 0.04 and 0.2 through float32 ops: within one ulp of 2u - 1
    assert v[0] == 0.0
    assert abs(v[6] - (-0.92)) <= np.spacing(np.float32(0.92)) and abs(v[5] - (-0.6)) <= np.spacing(np.float32(0.6))
    o.init_from_const(dif=0.9, spe=0.3, rgh=0.7)
    u = (o.textures.detach()[0, :, 3, 5].double().numpy() + 1) / 2
    np.testing.assert_allclose(u[[0, 1, 2, 5, 6, 7, 8]], [0.9, 0.9, 0.9, 0.7, 0.3, 0.3, 0.3], atol=1e-7)


def test_init_from_randn_statistics(pkg):
    """src/svbrdf.py:32-39: normal xy ~ clamp(N(0,1)/4, -1, 1); diffuse, specular
    and roughness u ~ clamp(N(0,1)/s + mean, 0, 1) with s = 8, 32, 16, stored as
    2u - 1.  Checked through the moments of the clamped normal (the clamp
    truncates specular at u = 0 for N < -1.28, ~10 % of its texels, which
    gives exact -1 values), and determinism under torch.manual_seed."""
    from scipy.stats import norm
    R = 256
    o = _optim_on_host(pkg, R)
    torch.manual_seed(1234)
    o.init_from_randn()
    x = o.textures.detach()[0].double().numpy()
    assert o.textures.requires_grad and x.min() >= -1 and x.max() <= 1
    n = x[3:5]
    assert abs(n.mean()) < 0.005 and abs(n.std() - 0.25) < 0.005
    for ch, mean, s in ((slice(0, 3), 0.5, 8), (slice(6, 9), 0.04, 32), (slice(5, 6), 0.2, 16)):
        u = (x[ch] + 1) / 2
        a = -mean * s                      # N below a clamps to u = 0
        p0 = norm.cdf(a)
        # mean of clamp(N/s + mean, 0, 1) (the upper clamp is > 6 sigma away)
        m = (mean * (1 - p0) + norm.pdf(a) / s)
        assert abs(u.mean() - m) < 3e-3 / s * 8, (ch, u.mean(), m)
        assert abs((u == 0).mean() - p0) < 0.01, (ch, (u == 0).mean(), p0)
    assert (x[6:9] == -1.0).mean() > 0.09             # exact -1 specular texels
    torch.manual_seed(1234)
    o2 = _optim_on_host(pkg, R)
    o2.init_from_randn()
    assert torch.equal(o.textures, o2.textures)


def test_material_recipe_is_uniform_per_channel(pkg):
    """albedo="material": one draw per channel for the whole map (diffuse,
    specular, roughness constant over the image, in their ranges); the
    normals are the per-texel recipe's, from the same seed."""
    syn = pkg.synthetic
    m = syn.ground_truth_textures(64, seed=4, albedo="material")
    t = syn.ground_truth_textures(64, seed=4)
    for c in (0, 1, 2, 5, 6, 7, 8):
        assert np.unique(m[c]).size == 1, c
    u = (m[:, 0, 0].astype(np.float64) + 1) / 2
    assert np.all((0.05 <= u[0:3]) & (u[0:3] <= 0.9)) and np.all((0.02 <= u[6:9]) & (u[6:9] <= 0.3))
    assert 0.1 <= u[5] <= 0.9
    np.testing.assert_array_equal(m[3:5], t[3:5])
    with pytest.raises(ValueError):
        syn.ground_truth_textures(8, albedo="pixel")


def test_gram_function_against_autograd(pkg):
    """compose.gram (per-image products, hand-written backward (dG + dG^T) F
    / (c h w)) against torch autograd of the bmm formula in FP64 on the CPU,
    and torch.autograd.gradcheck of it."""
    from svbrdf_amd import compose
    from tests import convnet_ref as R
    g = torch.Generator().manual_seed(9)
    f = torch.randn(3, 5, 4, 6, generator=g, dtype=torch.float64)
    probe = torch.randn(3, 5, 5, generator=g, dtype=torch.float64)
    a = f.clone().requires_grad_(True)
    b = f.clone().requires_grad_(True)
    ga, gb = compose.gram(a), R.gram_ref(b)
    torch.testing.assert_close(ga, gb, rtol=1e-12, atol=1e-14)
    (ga * probe).sum().backward()
    (gb * probe).sum().backward()
    torch.testing.assert_close(a.grad, b.grad, rtol=1e-12, atol=1e-14)
    assert torch.autograd.gradcheck(compose.gram, (f.clone().requires_grad_(True),))
