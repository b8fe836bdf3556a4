"""svbrdf_render_backward_noalloc: the render adjoint with its light-power
partial sums in the library's own device array (no workspace argument, no
allocation) against svbrdf_render_backward with a caller workspace, through
the C ABI: bit-identical outputs on two calls in a row, the second with a
smaller input (the array keeps the first call's contents)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _inputs(pkg, res, V, seed):
    syn = pkg.synthetic
    tex = syn.init_textures(syn.ground_truth_textures(res, seed=seed), seed=seed + 1)
    rng = np.random.default_rng(seed)
    cam, light = syn.flash_grid(jitter=3.0, rng=rng, n_views=V)
    gout = rng.standard_normal((V, 3, res, res)).astype(np.float32) * 1e-3
    t = lambda a: torch.tensor(a, device=DEV).contiguous()  # noqa: E731
    return t(tex), t(cam), t(light), t(syn.light_power()), t(gout), syn.PLANE_SIZE_CM


def _both(pkg, res, V, seed):
    L = pkg._lib
    tex, cam, light, P, gout, size = _inputs(pkg, res, V, seed)
    s = L.stream_handle(DEV)
    gt_ws = torch.full((9, res, res), 7.0, device=DEV)
    gp_ws = torch.full((3,), 7.0, device=DEV)
    nbytes = L.query("svbrdf_render_backward_workspace_bytes", res)
    ws = torch.zeros(nbytes // 4, device=DEV)
    L.call("svbrdf_render_backward", L.ptr(tex), res, V, size, L.ptr(cam), L.ptr(light), L.ptr(P), L.ptr(gout),
           L.ptr(gt_ws), L.ptr(gp_ws), L.ptr(ws), nbytes, s)
    gt_na = torch.full((9, res, res), -7.0, device=DEV)
    gp_na = torch.full((3,), -7.0, device=DEV)
    L.call("svbrdf_render_backward_noalloc", L.ptr(tex), res, V, size, L.ptr(cam), L.ptr(light), L.ptr(P),
           L.ptr(gout), L.ptr(gt_na), L.ptr(gp_na), s)
    torch.cuda.synchronize()
    return (gt_ws.cpu().numpy(), gp_ws.cpu().numpy()), (gt_na.cpu().numpy(), gp_na.cpu().numpy())


def test_noalloc_adjoint_matches_workspace_entry_point(pkg):
    for res, V, seed in ((512, 9, 1), (100, 5, 2)):
        print(f"noalloc adjoint res={res} V={V}", flush=True)
        (gt_a, gp_a), (gt_b, gp_b) = _both(pkg, res, V, seed)
        assert np.all(np.isfinite(gt_a)) and np.all(np.isfinite(gp_a)) and np.abs(gp_a).sum() > 0
        np.testing.assert_array_equal(gt_b.view(np.uint32), gt_a.view(np.uint32))
        np.testing.assert_array_equal(gp_b.view(np.uint32), gp_a.view(np.uint32))


def test_noalloc_adjoint_without_light_gradient_and_limits(pkg):
    L = pkg._lib
    res, V = 64, 9
    tex, cam, light, P, gout, size = _inputs(pkg, res, V, 3)
    s = L.stream_handle(DEV)
    a = torch.empty((9, res, res), device=DEV)
    b = torch.empty((9, res, res), device=DEV)
    L.call("svbrdf_render_backward", L.ptr(tex), res, V, size, L.ptr(cam), L.ptr(light), L.ptr(P), L.ptr(gout),
           L.ptr(a), None, None, 0, s)
    L.call("svbrdf_render_backward_noalloc", L.ptr(tex), res, V, size, L.ptr(cam), L.ptr(light), L.ptr(P),
           L.ptr(gout), L.ptr(b), None, s)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    with pytest.raises(ValueError, match="4096"):
        L.call("svbrdf_render_backward_noalloc", L.ptr(tex), 4097, V, size, L.ptr(cam), L.ptr(light), L.ptr(P),
               L.ptr(gout), L.ptr(b), None, s)
