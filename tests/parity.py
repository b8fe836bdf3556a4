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
  Pixels at a kink of the output clamp may be left out of the norms: a
  pixel one of whose pre-clamp image values (FP64) lies within 2e-6
  relative of a clamp bound (1e-6 or 1).  There rounding decides whether a
  whole view's contribution passes the clamp's mask, in any FP32
  implementation (the reference's included), and the gradient jumps.  At
  most 1e-4 of the pixels may be left out.
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


KINK_WINDOW = 2e-6
MAX_KINK_FRACTION = 1e-4


def clamp_kinks(img64, eps=1e-6):
    """[R, R] mask of the pixels whose pre-clamp image ([V, 3, R, R], FP64)
    lies within KINK_WINDOW relative of a clamp bound in some view and
    channel."""
    img64 = np.asarray(img64)
    near = (np.abs(img64 / eps - 1.0) <= KINK_WINDOW) | (np.abs(img64 - 1.0) <= KINK_WINDOW)
    return near.any(axis=(0, 1))


def check_gradient(g, g32, g64, label="", exclude=None):
    g = np.asarray(g, np.float64).reshape(g64.shape)
    if exclude is not None and exclude.any():
        assert exclude.mean() <= MAX_KINK_FRACTION, f"{label}: {int(exclude.sum())} kink pixels"
        keep = ~exclude
        g, g32, g64 = g * keep, np.asarray(g32) * keep, np.asarray(g64) * keep
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
