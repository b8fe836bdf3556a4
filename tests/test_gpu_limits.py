"""The largest resolution the library accepts (svbrdf_max_res(), 7680): every
texture plane then sits at a byte offset up to 8 * 7680^2 * 4 within one
buffer range.  The last rows of the image -- the largest offsets of every
plane -- are checked against the oracle: the forward render, and the gradient
the fused step used (recovered from Adam's first moment, m1 = 0.1 g).
Criteria: tests/parity.py."""
import numpy as np
import pytest
import torch

from oracle import microfacet_np as M
from tests import parity

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def test_forward_and_fused_step_at_max_resolution(pkg):
    res = pkg._lib.load().svbrdf_max_res()
    assert res == 7680
    print(f"max-res test at {res}^2", flush=True)
    V, size = 1, 17.0
    rng = np.random.default_rng(0)
    tex = rng.uniform(-0.9, 0.9, (9, res, res)).astype(np.float32)
    tex[3:5] *= np.float32(0.3)
    cam = np.array([[2.0, -1.0, 30.0]], np.float32)
    light = cam.copy()
    P = np.full(3, 1500.0, np.float32)
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    r = pkg.Microfacet(res, V, size, cl, DEV)
    t = torch.from_numpy(tex).to(DEV)[None].contiguous()
    band = slice(res - 4, res)
    with torch.no_grad():
        out = r.eval(t)
        target = r.eval(torch.clamp(t + 0.05, -1, 1))
    out_band = out[:, :, band].cpu().numpy()
    tgt_band = target[:, :, band].cpu().numpy()
    del out
    o32 = M.render(tex, res, size, cam, light, P, np.float32, rows=band)
    o64 = M.render(tex, res, size, cam, light, P, np.float64, rows=band)
    sig = M.fp32_noise_scale(tex, res, size, cam, light, P, rows=band)
    parity.check_forward(out_band, o32, o64, sig, "max-res band")

    fs = pkg.FusedStep(r, t, target, 0.01, loss="log_l2")
    del target
    fs.step()
    torch.cuda.synchronize()
    g_gpu = fs.exp_avg[0][:, band].cpu().numpy().astype(np.float64) / np.float64(np.float32(0.1))
    assert np.isfinite(fs.loss())
    N = V * 3 * res * res
    grads = {}
    for dt in (np.float32, np.float64):
        ob = M.render(tex, res, size, cam, light, P, dt, rows=band)
        y = np.clip(tgt_band.astype(dt), M.target_floor(dt), dt(1))
        gout = (dt(2) * (np.log(ob) - np.log(y)) / dt(N) / ob).astype(dt)
        grads[dt], _ = M.render_vjp(tex, res, size, cam, light, P, gout, dt, rows=band)
    parity.check_gradient(g_gpu, grads[np.float32], grads[np.float64], "max-res step band")
