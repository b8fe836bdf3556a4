"""Seeded synthetic SVBRDF maps and flash geometry.

The reference ships no data (its photos and checkpoints are external), so
every workload here is generated.  The recipe:

* normal: white noise N(0,1) on an R x R torus, Gaussian-filtered with
  sigma = sigma_px (default 4 px at 256^2, scaled by R/256), rescaled so
  that max|h| = amplitude (0.5); n = normalize(-dh/dx, -dh/dy, 1) with
  central differences in pixel units; (nx, ny) go to channels 3:5 as in
  src/svbrdf.py:160,166 (normals are stored directly, not as 2u-1).
* diffuse u ~ U[0.05, 0.9] per RGB, specular u ~ U[0.02, 0.3] per RGB,
  roughness u ~ U[0.1, 0.9]; stored as t = 2u - 1 (src/svbrdf.py:161-163).
  Two readings of "per RGB": ``albedo="texel"`` draws every texel
  independently (each map covers its whole value range in one image, which
  the parity tests want), ``albedo="material"`` draws one value per channel
  for the whole map (a uniformly coloured material whose spatial detail is
  the smooth normal field; a coarser grid represents it, which
  coarse-to-fine stage chaining presupposes).  The benchmark workloads use
  "material".
* views: a 3x3 grid (x, y) in {-8, 0, 8}^2 cm at z = 30 cm, +y row first
  and x from - to + within a row (the order of data/random/render.json:7-24),
  light co-located with the camera, light power [1500, 1500, 1500];
  optional per-view jitter U[-j, j] cm on x and y.
* plane size 17 cm.
* optimisation start: ground truth + N(0, 0.1), clamped to [-1, 1].

Seeds: 0 ground truth, 1 init noise, 3 batch sizes/views/jitter.
"""
import numpy as np

PLANE_SIZE_CM = 17.0
LIGHT_POWER = 1500.0


def _gaussian_filter_periodic(a, sigma):
    r = a.shape[0]
    f = np.fft.fftfreq(r)
    g = np.exp(-2.0 * (np.pi * sigma) ** 2 * (f[:, None] ** 2 + f[None, :] ** 2))
    return np.real(np.fft.ifft2(np.fft.fft2(a) * g))


def height_field_normals(res, rng, sigma_px=None, amplitude=0.5):
    """(nx, ny) of a Gaussian-filtered white-noise height field, float32."""
    if sigma_px is None:
        sigma_px = 4.0 * res / 256.0
    if res < 2:
        return np.zeros((res, res), np.float32), np.zeros((res, res), np.float32)
    h = _gaussian_filter_periodic(rng.standard_normal((res, res)), sigma_px)
    h *= amplitude / max(np.abs(h).max(), 1e-30)
    dy, dx = np.gradient(h)
    # image rows run toward -y (row 0 is at +y), so dh/dy_world = -dh/drow
    gx, gy = dx, -dy
    n = np.stack((-gx, -gy, np.ones_like(h)), 0)
    n /= np.linalg.norm(n, axis=0, keepdims=True)
    return n[0].astype(np.float32), n[1].astype(np.float32)


def ground_truth_textures(res, seed=0, sigma_px=None, albedo="texel"):
    """``[9, R, R]`` float32 maps in [-1, 1]: dif 0:3, normal 3:5, rgh 5, spe 6:9.
    albedo = "texel" (an independent draw per texel) or "material" (one draw
    per channel for the whole map)."""
    if albedo not in ("texel", "material"):
        raise ValueError(f"albedo must be 'texel' or 'material', got {albedo!r}")
    rng = np.random.Generator(np.random.PCG64(seed))
    nx, ny = height_field_normals(res, rng, sigma_px)
    n = (res, res) if albedo == "texel" else (1, 1)
    dif = rng.uniform(0.05, 0.9, (3, *n))
    spe = rng.uniform(0.02, 0.3, (3, *n))
    rgh = rng.uniform(0.1, 0.9, (1, *n))
    t = np.empty((9, res, res), np.float32)
    t[0:3] = dif * 2 - 1
    t[3] = nx
    t[4] = ny
    t[5:6] = rgh * 2 - 1
    t[6:9] = spe * 2 - 1
    return t


def init_textures(gt, seed=1, std=0.1):
    """Optimisation start: gt + N(0, std), clamped to [-1, 1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.clip(gt + rng.normal(0.0, std, gt.shape), -1, 1).astype(np.float32)


def flash_grid(offset=8.0, height=30.0, jitter=0.0, rng=None, n_views=9):
    """Co-located camera/flash positions ``(cam [V,3], light [V,3])`` in cm.

    The first 9 views walk the 3x3 grid (+y row first); further views (for
    V > 9) repeat the grid.  ``jitter`` adds U[-j, j] to x and y of each view.
    """
    base = np.array([[x, y, height] for y in (offset, 0.0, -offset)
                     for x in (-offset, 0.0, offset)], np.float64)
    idx = np.arange(n_views) % 9
    cam = base[idx].copy()
    if jitter:
        cam[:, :2] += rng.uniform(-jitter, jitter, (n_views, 2))
    cam = cam.astype(np.float32)
    return cam, cam.copy()


def light_power():
    return np.full(3, LIGHT_POWER, np.float32)


def downsample(tex, factor):
    """Area (box) average of ``[9, R, R]`` maps by an integer factor: the same
    material at a lower resolution (used to build coarse-to-fine stages)."""
    c, r, _ = tex.shape
    if r % factor:
        raise ValueError(f"resolution {r} is not divisible by {factor}")
    return tex.reshape(c, r // factor, factor, r // factor, factor).mean(axis=(2, 4)).astype(np.float32)
