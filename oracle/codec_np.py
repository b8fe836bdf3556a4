"""NumPy restatement of the map / target codec around the optimisation path.
TEST INFRASTRUCTURE ONLY (see oracle/__init__).

What the reference does between PNG files and the tensors the renderer
takes (src/imageio.py:11-76, src/svbrdf.py:126-132,150-221):

* ``imread(fn, flag, (res, res))`` (imageio.py:11-51): cv2 reads 8-bit BGR,
  ``imresize`` resizes the uint8 image with ``cv2.INTER_LANCZOS4``
  (imageio.py:75-76), the result becomes float32 ``u / 255``, then per flag:
  "srgb" reorders to RGB, "rough" averages the 3 channels (or keeps 1),
  "normal" reorders to RGB, maps ``*2 - 1`` and divides by the per-pixel
  2-norm (np.linalg.norm: sqrt of the sequential sum of squares).
* ``load_textures_th`` (svbrdf.py:150-166) maps diffuse / specular /
  roughness through ``*2 - 1`` and packs [dif 0:3, normal xy 3:5, rgh 5,
  spe 6:9]; ``load_images_th`` (svbrdf.py:191-204) stacks "srgb" images.
* ``save_textures_th`` (svbrdf.py:168-189) writes ``(t + 1) / 2`` for
  diffuse / specular / roughness and ``reconstruct_normal`` (svbrdf.py:
  126-132: clamp to [-1, 1], z = sqrt(1 - clamp(x^2 + y^2, 0, 1)), divide by
  the 2-norm) for the normal; ``imwrite`` (imageio.py:54-72) clips ("srgb",
  "rough": [0, 1]; "normal": [-1, 1] then ``(n + 1) / 2``) and stores
  ``(v * 255).astype(uint8)``, i.e. truncation toward zero.
  ``save_images_th`` (svbrdf.py:206-221) writes images through "srgb".

Arrays here are RGB-ordered (the channel reorders of the reference cancel
between cv2's BGR and the file's RGB).  Every float op is float32 and
rounds once, in the reference's order, with no fused multiply-add.

The resize restates OpenCV's INTER_LANCZOS4 for 8-bit images as published
in its source (modules/imgproc/src/resize.cpp): destination pixel d maps to
source coordinate f = (d + 0.5) * scale - 0.5 with scale = 1 / (dst / src)
in double, rounded to float; taps floor(f) - 3 .. floor(f) + 4 with the
border replicated; weights are the Lanczos-4 window sinc(t) sinc(t / 4) at
t = frac + 3 - i, normalised to sum 1 in float, then scaled by 2^11 and
rounded half-to-even to int16; horizontal sums are exact integers, vertical
sums of those are shifted right by 22 with rounding and saturated to
[0, 255].  Same-size resizes are copies (cv2.resize returns the source).
cv2 itself is not installed anywhere this project runs, so bit parity with
cv2 is "parity unpinned": the restatement is pinned by identity at equal
size, by an independent FP64 brute-force Lanczos-4 (within one level), by
the weight invariants (partition of unity, mirror symmetry, the delta at
integer offsets) and by flat and separable inputs (tests/test_codec.py).
"""
import math

import numpy as np

COEF_BITS = 11                      # fixed-point weight bits of the 8-bit resize
COEF_SCALE = 1 << COEF_BITS
_S45 = math.sqrt(0.5)
# (-1)^(i+1) (cos(i pi/4), sin(i pi/4)): the numerator of tap i is
# A_i sin(-a) + B_i cos(-a) = (-1)^i sin(a - i pi/4), a = (x + 3) pi / 4
_AB = [(-1.0, 0.0), (_S45, _S45), (0.0, -1.0), (-_S45, _S45),
       (1.0, 0.0), (-_S45, -_S45), (0.0, 1.0), (_S45, -_S45)]


def lanczos4_weights(x):
    """Eight float32 weights for fractional offset x (float32 in [0, 1)).

    The Lanczos-4 kernel sinc(t) sinc(t/4) at t_i = x + 3 - i equals
    (4 / pi^2 t_i^2) sin(pi t_i) sin(pi t_i / 4); sin(pi t_i) = (-1)^i
    sin(4a) is common to all taps and cancels in the normalisation, leaving
    (-1)^i sin(a - i pi/4) / t_i^2 up to a common factor (here with
    (pi t_i / 4)^2 as the denominator, in double).  A tap at t = 0 gets
    weight 1e30, which normalises to the unit impulse."""
    x = np.float32(x)
    xp3 = np.float32(x + np.float32(3))
    a0 = -float(xp3) * math.pi * 0.25
    s0, c0 = math.sin(a0), math.cos(a0)
    w = np.zeros(8, np.float32)
    total = np.float32(0)
    for i in range(8):
        t = np.float32(xp3 - np.float32(i))
        if abs(float(t)) >= 1e-6:
            y = -float(t) * math.pi * 0.25
            A, B = _AB[i]
            w[i] = np.float32((A * s0 + B * c0) / (y * y))
        else:
            w[i] = np.float32(1e30)
        total = np.float32(total + w[i])
    inv = np.float32(np.float32(1) / total)
    return (w * inv).astype(np.float32)


def _fixed_weights(x):
    w = lanczos4_weights(x) * np.float32(COEF_SCALE)
    return np.clip(np.rint(w.astype(np.float64)), -32768, 32767).astype(np.int64)   # rint = half to even


def axis_plan(n_src, n_dst):
    """Per destination index: first tap (floor(f) - 3) and int16 weights."""
    scale = 1.0 / (float(n_dst) / float(n_src))
    first = np.empty(n_dst, np.int64)
    wts = np.empty((n_dst, 8), np.int64)
    for d in range(n_dst):
        f = np.float32((d + 0.5) * scale - 0.5)
        s = math.floor(float(f))
        frac = np.float32(f - np.float32(s))
        first[d] = s - 3
        wts[d] = _fixed_weights(frac)
    return first, wts


def resize_lanczos4_u8(src, dst_h, dst_w):
    """cv2.resize(src, (dst_w, dst_h), interpolation=INTER_LANCZOS4) for a
    uint8 [H, W] or [H, W, C] image (src/imageio.py:75-76)."""
    src = np.asarray(src)
    assert src.dtype == np.uint8
    squeeze = src.ndim == 2
    if squeeze:
        src = src[:, :, None]
    h, w, _ = src.shape
    if (h, w) == (dst_h, dst_w):
        out = src.copy()
        return out[:, :, 0] if squeeze else out
    fy, wy = axis_plan(h, dst_h)
    fx, wx = axis_plan(w, dst_w)
    s = src.astype(np.int64)
    cols = np.clip(fx[:, None] + np.arange(8)[None, :], 0, w - 1)           # [dw, 8]
    horiz = np.einsum("hdkc,dk->hdc", s[:, cols, :], wx)                     # exact int
    rows = np.clip(fy[:, None] + np.arange(8)[None, :], 0, h - 1)           # [dh, 8]
    vert = np.einsum("dkwc,dk->dwc", horiz[rows, :, :], wy)
    out = np.clip((vert + (1 << (2 * COEF_BITS - 1))) >> (2 * COEF_BITS), 0, 255).astype(np.uint8)
    return out[:, :, 0] if squeeze else out


def u8_to_unit(u):
    """im.astype("float32") / 255 (src/imageio.py:17-18)."""
    return (np.asarray(u).astype(np.float32) / np.float32(255)).astype(np.float32)


def imread_u8(u, flag, dim=None):
    """src/imageio.py:11-51 after the file decode, for an RGB-ordered uint8
    image; dim = (w, h) as cv2 takes it."""
    u = np.asarray(u, np.uint8)
    if dim is not None:
        u = resize_lanczos4_u8(u, dim[1], dim[0])
    im = u8_to_unit(u)
    if flag == "srgb":
        if im.ndim != 3 or im.shape[2] != 3:
            raise ValueError("srgb image must have 3 channels")
        return im
    if flag == "rough":
        if im.ndim == 3 and im.shape[2] == 3:
            s = np.float32(im[:, :, 0] + im[:, :, 1])
            s = (s + im[:, :, 2]).astype(np.float32)
            return (s / np.float32(3)).astype(np.float32)
        if im.ndim == 2:
            return im
        raise ValueError("rough image must have 1 or 3 channels")
    if flag == "normal":
        if im.ndim != 3 or im.shape[2] != 3:
            raise ValueError("normal image must have 3 channels")
        n = (im * np.float32(2) - np.float32(1)).astype(np.float32)
        sq = (n * n).astype(np.float32)
        ss = ((sq[:, :, 0] + sq[:, :, 1]).astype(np.float32) + sq[:, :, 2]).astype(np.float32)
        norm = np.sqrt(ss).astype(np.float32)
        return (n / norm[:, :, None]).astype(np.float32)
    return im


def load_textures(dif_u8, nom_u8, spe_u8, rgh_u8, res):
    """SvbrdfIO.load_textures_th (src/svbrdf.py:150-166) -> [9, res, res]."""
    dim = (res, res)
    normal = imread_u8(nom_u8, "normal", dim)
    diffuse = imread_u8(dif_u8, "srgb", dim)
    specular = imread_u8(spe_u8, "srgb", dim)
    roughness = imread_u8(rgh_u8, "rough", dim)
    two, one = np.float32(2), np.float32(1)
    out = np.empty((9, res, res), np.float32)
    out[0:3] = (diffuse * two - one).transpose(2, 0, 1)
    out[3:5] = normal[:, :, :2].transpose(2, 0, 1)
    out[5] = roughness * two - one
    out[6:9] = (specular * two - one).transpose(2, 0, 1)
    return out


def load_images(images_u8, res):
    """SvbrdfIO.load_images_th (src/svbrdf.py:191-204): [n, H, W, 3] uint8 ->
    [n, 3, res, res] float32."""
    return np.stack([imread_u8(im, "srgb", (res, res)).transpose(2, 0, 1) for im in images_u8]).astype(np.float32)


def to_u8(v):
    """(v * 255).astype("uint8") for v already clipped to [0, 1] (src/imageio.py:71)."""
    return (np.asarray(v, np.float32) * np.float32(255)).astype(np.float32).astype(np.uint8)


def reconstruct_normal_io(nxy):
    """SvbrdfIO.reconstruct_normal (src/svbrdf.py:126-132) on [2, R, R]:
    x^2 + y^2 + z^2 summed in that order."""
    x = np.clip(nxy[0], -1, 1).astype(np.float32)
    y = np.clip(nxy[1], -1, 1).astype(np.float32)
    xy = np.clip((x * x + y * y).astype(np.float32), 0, 1).astype(np.float32)
    z = np.sqrt((np.float32(1) - xy).astype(np.float32)).astype(np.float32)
    ss = ((x * x + y * y).astype(np.float32) + z * z).astype(np.float32)
    norm = np.sqrt(ss).astype(np.float32)
    return np.stack([x / norm, y / norm, z / norm]).astype(np.float32)


def save_textures(tex):
    """SvbrdfIO.save_textures_th (src/svbrdf.py:168-189) + imwrite flags
    (src/imageio.py:54-72) on [9, R, R] float32 -> (dif [R,R,3], nom [R,R,3],
    spe [R,R,3], rgh [R,R]) uint8, RGB order."""
    tex = np.asarray(tex, np.float32)
    one, two = np.float32(1), np.float32(2)
    half = lambda t: ((t + one).astype(np.float32) / two).astype(np.float32)
    dif = to_u8(np.clip(half(tex[0:3]), 0, 1)).transpose(1, 2, 0)
    spe = to_u8(np.clip(half(tex[6:9]), 0, 1)).transpose(1, 2, 0)
    rgh = to_u8(np.clip(half(tex[5]), 0, 1))
    n = reconstruct_normal_io(tex[3:5])
    nom = to_u8(half(np.clip(n, -1, 1))).transpose(1, 2, 0)
    return (np.ascontiguousarray(dif), np.ascontiguousarray(nom), np.ascontiguousarray(spe),
            np.ascontiguousarray(rgh))


def save_images(images):
    """SvbrdfIO.save_images_th pixel part: [n, 3, R, R] float32 -> [n, R, R, 3]
    uint8 through imwrite "srgb"."""
    images = np.asarray(images, np.float32)
    return np.ascontiguousarray(to_u8(np.clip(images, 0, 1)).transpose(0, 2, 3, 1))


def lanczos4_fp64(src, dst_h, dst_w):
    """Independent check: the Lanczos-4 resample in float64 from the kernel
    definition (sinc(t) sinc(t/4), normalised per axis), replicated border,
    rounded to nearest.  Differs from the fixed-point resize by at most one
    level on ordinary images."""
    src = np.asarray(src, np.float64)
    squeeze = src.ndim == 2
    if squeeze:
        src = src[:, :, None]

    def plan(n_src, n_dst):
        scale = n_src / n_dst
        idx = np.empty((n_dst, 8), np.int64)
        w = np.empty((n_dst, 8))
        for d in range(n_dst):
            f = (d + 0.5) * scale - 0.5
            s = math.floor(f)
            fr = f - s
            t = fr + 3 - np.arange(8)
            k = np.sinc(t) * np.sinc(t / 4)
            w[d] = k / k.sum()
            idx[d] = np.clip(s - 3 + np.arange(8), 0, n_src - 1)
        return idx, w

    h, wd, _ = src.shape
    iy, wy = plan(h, dst_h)
    ix, wx = plan(wd, dst_w)
    horiz = np.einsum("hdkc,dk->hdc", src[:, ix, :], wx)
    vert = np.einsum("dkwc,dk->dwc", horiz[iy, :, :], wy)
    out = np.clip(np.rint(vert), 0, 255)
    return out[:, :, 0] if squeeze else out
