"""Parity criteria shared by the GPU tests (tolerances are stated here).

Forward images (FP32):
  every element: |gpu - ref32| <= 1e-6 + 1e-5 |ref32| + 8 sigma,
  where sigma is the FP64 response of the output to a one-ulp change of the
  ill-conditioned dot products n.v, n.l, n.h (oracle fp32_noise_scale); and
  the GPU's max relative error against the FP64 render is at most twice the
  reference FP32's own (+1e-6).  On well-conditioned pixels sigma ~ 0 and the
  bound is the plain 1e-5 relative / 1e-6 absolute; additionally at least
  99.9 % of all elements must meet that plain bound against the FP64 render
  (or, where the reference FP32 itself meets it for fewer, no fewer than the
  reference FP32 does, less 0.1 %).

Gradients (FP32), per texture channel:
  ||g_gpu - g64|| <= max(1e-4 ||g64||, 2 ||g_ref32 - g64||): within 1e-4
  relative of the FP64 gradient, or no worse than twice the error of the
  reference's own FP32 autograd gradient where FP32 cannot reach 1e-4; and
  ||g_gpu - g_ref32|| <= 1e-4 ||g_ref32|| + 3 ||g_ref32 - g64||.
"""
import numpy as np


def check_forward(out, ref32, ref64, sigma, label=""):
    out = np.asarray(out, np.float64)
    err = np.abs(out - ref32)
    plain = 1e-6 + 1e-5 * np.abs(ref32)
    bound = plain + 8 * sigma
    bad = err > bound
    assert not bad.any(), (f"{label}: {int(bad.sum())} / {bad.size} elements out of tolerance; worst excess "
                           f"{(err - bound).max():.3e} at {np.unravel_index((err - bound).argmax(), err.shape)}")
    frac_plain = float(np.mean(err <= plain))
    frac64 = float(np.mean(np.abs(out - ref64) <= 1e-6 + 1e-5 * np.abs(ref64)))
    # where FP32 itself cannot reach 1e-5 of FP64 (e.g. the fp32 constant 1 - 1e-6
    # of src/microfacet.py:67 at grazing normals) the reference FP32 sets the bar
    frac64_ref = float(np.mean(np.abs(ref32 - ref64) <= 1e-6 + 1e-5 * np.abs(ref64)))
    assert frac64 >= min(0.999, frac64_ref - 1e-3), (
        f"{label}: only {frac64:.5f} of elements within 1e-5 rel / 1e-6 abs of FP64 (reference FP32: {frac64_ref:.5f})")
    e_gpu = float((np.abs(out - ref64) / ref64).max())
    e_ref = float((np.abs(ref32 - ref64) / ref64).max())
    assert e_gpu <= 2 * e_ref + 1e-6, f"{label}: max rel err vs FP64 {e_gpu:.3e} > 2 x reference FP32's {e_ref:.3e}"
    return dict(max_rel_vs_fp64=e_gpu, ref32_max_rel_vs_fp64=e_ref, frac_plain_vs_ref32=frac_plain,
                frac_plain_vs_fp64=frac64)


def check_gradient(g, g32, g64, label=""):
    g = np.asarray(g, np.float64).reshape(g64.shape)
    out = []
    for c in range(g64.shape[0]):
        n64 = np.linalg.norm(g64[c])
        n32 = np.linalg.norm(g32[c])
        e64 = np.linalg.norm(g[c] - g64[c])
        eref = np.linalg.norm(g32[c] - g64[c])
        assert e64 <= max(1e-4 * n64, 2 * eref) + 1e-30, (
            f"{label} ch{c}: |g-g64|/|g64| = {e64 / max(n64, 1e-300):.3e}, reference FP32 {eref / max(n64, 1e-300):.3e}")
        e32 = np.linalg.norm(g[c] - g32[c])
        assert e32 <= 1e-4 * n32 + 3 * eref + 1e-30, (
            f"{label} ch{c}: |g-g32|/|g32| = {e32 / max(n32, 1e-300):.3e}")
        out.append(e64 / max(n64, 1e-300))
    return out
