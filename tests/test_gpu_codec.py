"""Map / target codec on the GPU against the oracle (oracle/codec_np.py):
byte outputs bit for bit, float outputs bit for bit (the kernels follow the
oracle's float32 op order with no fused multiply-adds and correctly rounded
division / square root), through the C ABI and through SvbrdfIO with real
PNG files."""
import json

import numpy as np
import pytest
import torch

from oracle import codec_np as C

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def smooth_u8(rng, shape):
    yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
    base = 127 + 100 * np.sin(xx / 6.0 + rng.uniform(0, 3)) * np.cos(yy / 4.0)
    if len(shape) == 3:
        base = np.repeat(base[:, :, None], shape[2], 2) + rng.uniform(-20, 20, (1, 1, shape[2]))
    return np.clip(base + rng.normal(0, 25, base.shape), 0, 255).astype(np.uint8)


@pytest.mark.parametrize("shape,dst", [((64, 64, 3), (256, 256)), ((256, 256, 3), (64, 64)), ((50, 31, 1), (77, 64)),
                                       ((33, 47, 3), (20, 90)), ((128, 96, 4), (200, 150)), ((37, 53, 3), (37, 53)),
                                       ((1, 1, 3), (5, 7)), ((300, 200, 2), (150, 401)), ((512, 512, 3), (1024, 1024))])
def test_resize_matches_oracle_bit_for_bit(pkg, shape, dst):
    from svbrdf_amd import imageio as io
    print(f"resize {shape} -> {dst}", flush=True)
    rng = np.random.default_rng(sum(shape) + sum(dst))
    im = smooth_u8(rng, shape)
    out = io.imresize(torch.from_numpy(im).to(DEV), (dst[1], dst[0])).cpu().numpy()
    ref = C.resize_lanczos4_u8(im, *dst)
    assert out.shape == ref.shape
    np.testing.assert_array_equal(out, ref)


def on_device(a, misalign=False):
    """A device copy; misalign=True places it one byte (or float) past an
    aligned address, so that the library takes its one-pixel-per-thread path
    (four pixels per thread need 4-byte aligned bytes and 16-byte aligned
    floats and a pixel count divisible by 4)."""
    t = torch.from_numpy(np.ascontiguousarray(a))
    if not misalign:
        return t.to(DEV)
    buf = torch.empty(t.numel() + 1, dtype=t.dtype, device=DEV)
    v = buf[1:].view(t.shape)
    v.copy_(t)
    return v


# (17, 29) and the misaligned copies: one pixel per thread; (16, 20): four
@pytest.mark.parametrize("hw,misalign", [((17, 29), False), ((16, 20), False), ((16, 20), True)])
@pytest.mark.parametrize("flag,channels", [(None, 3), (None, 1), ("srgb", 3), ("rough", 3), ("rough", 1),
                                           ("normal", 3), (None, 4)])
def test_decode_flags_match_oracle(pkg, flag, channels, hw, misalign):
    from svbrdf_amd import imageio as io
    rng = np.random.default_rng(channels)
    ims = rng.integers(0, 256, (3, *hw, channels), dtype=np.uint8)
    out = io.decode(on_device(ims, misalign), flag).cpu().numpy()
    for k in range(3):
        src = ims[k] if channels > 1 else ims[k, :, :, 0]
        ref = C.imread_u8(src, flag)
        ref = ref[None] if ref.ndim == 2 else ref.transpose(2, 0, 1)
        np.testing.assert_array_equal(out[k], ref)


@pytest.mark.parametrize("hw,misalign", [((19, 23), False), ((16, 12), False), ((16, 12), True)])
@pytest.mark.parametrize("flag,channels", [("srgb", 3), ("rough", 1), ("normal", 3), (None, 3)])
def test_encode_flags_match_oracle(pkg, flag, channels, hw, misalign):
    from svbrdf_amd import imageio as io
    rng = np.random.default_rng(10 + channels)
    x = rng.uniform(-1.3, 1.3, (2, channels, *hw)).astype(np.float32)
    x[0, 0, 0, :4] = [0.0, 1.0, 0.5, -0.0]
    out = io.encode(on_device(x, misalign), flag).cpu().numpy()
    for k in range(2):
        v = x[k].transpose(1, 2, 0)
        if flag == "normal":
            ref = C.to_u8(((np.clip(v, -1, 1) + np.float32(1)).astype(np.float32) / np.float32(2)).astype(np.float32))
        else:
            ref = C.to_u8(np.clip(v, 0, 1))
        np.testing.assert_array_equal(out[k], ref)


@pytest.mark.parametrize("res,rgh_channels,misalign", [(64, 1, False), (64, 3, False), (33, 3, False), (1, 1, False),
                                                        (64, 3, True)])
def test_textures_from_u8_matches_oracle(pkg, res, rgh_channels, misalign):
    from svbrdf_amd.svbrdf_io import textures_from_maps_u8
    rng = np.random.default_rng(res + rgh_channels)
    dif, nom, spe = (rng.integers(0, 256, (res, res, 3), dtype=np.uint8) for _ in range(3))
    rgh = rng.integers(0, 256, (res, res, 3) if rgh_channels == 3 else (res, res), dtype=np.uint8)
    t = textures_from_maps_u8(*(on_device(m, misalign) for m in (dif, nom, spe, rgh)), res)
    np.testing.assert_array_equal(t[0].cpu().numpy(), C.load_textures(dif, nom, spe, rgh, res))


@pytest.mark.parametrize("res,misalign", [(64, False), (33, False), (1, False), (64, True)])
def test_textures_to_u8_matches_oracle(pkg, res, misalign):
    from svbrdf_amd.svbrdf_io import maps_u8_from_textures
    rng = np.random.default_rng(100 + res)
    tex = rng.uniform(-1.2, 1.2, (9, res, res)).astype(np.float32)
    tex[3:5, 0, 0] = [0.0, 0.0]              # flat normal
    out = [m.cpu().numpy() for m in maps_u8_from_textures(on_device(tex, misalign)[None])]
    ref = C.save_textures(tex)
    for name, a, b in zip(("dif", "nom", "spe", "rgh"), out, ref):
        np.testing.assert_array_equal(a, b, err_msg=name)


def test_reload_textures_is_save_then_load(pkg):
    rng = np.random.default_rng(7)
    tex = np.clip(rng.normal(0, 0.5, (9, 64, 64)), -1, 1).astype(np.float32)
    got = pkg.reload_textures(torch.from_numpy(tex).to(DEV), 128)[0].cpu().numpy()
    np.testing.assert_array_equal(got, C.load_textures(*C.save_textures(tex), 128))


def test_svbrdf_io_files_end_to_end(pkg, tmp_path):
    from svbrdf_amd import imageio as io
    rng = np.random.default_rng(11)
    ref_dir, tgt_dir = tmp_path / "ref", tmp_path / "target"
    ref_dir.mkdir()
    tgt_dir.mkdir()
    maps = {"dif": smooth_u8(rng, (300, 300, 3)), "nom": smooth_u8(rng, (300, 300, 3)),
            "spe": smooth_u8(rng, (300, 300, 3)), "rgh": smooth_u8(rng, (300, 300))}
    for k, m in maps.items():
        io.write_png(ref_dir / f"{k}.png", m)
    idx = [0, 2, 5]
    views = {i: smooth_u8(rng, (200, 200, 3)) for i in idx}
    for i, v in views.items():
        io.write_png(tgt_dir / f"{i:02d}.png", v)
    cam = np.round(rng.uniform(-10, 10, (6, 3)), 3).tolist()
    (tmp_path / "cfg.json").write_text(json.dumps({
        "reference_dir": "ref", "target_dir": "target", "optimize_dir": "opt", "rerender_dir": "rr",
        "im_size": 17.0, "idx": idx, "camera_pos": cam, "light_pos": cam, "light_pow": [1500, 1500, 1500]}))
    sio = pkg.SvbrdfIO(tmp_path / "cfg.json", DEV)
    assert sio.n_of_imgs == 3 and sio.im_size == 17.0 and sio.target_dir == tgt_dir
    np.testing.assert_array_equal(sio.cl[0].cpu().numpy(), np.float32(cam)[idx])
    tex = sio.load_textures_th(sio.reference_dir, 256)
    assert tex.shape == (1, 9, 256, 256)
    np.testing.assert_array_equal(tex[0].cpu().numpy(),
                                  C.load_textures(maps["dif"], maps["nom"], maps["spe"], maps["rgh"], 256))
    imgs = sio.load_images_th(sio.target_dir, 128)
    np.testing.assert_array_equal(imgs.cpu().numpy(), C.load_images([views[i] for i in idx], 128))
    sio.save_textures_th(tex, sio.optimize_dir)
    for k, ref in zip(("dif", "nom", "spe", "rgh"), C.save_textures(tex[0].cpu().numpy())):
        np.testing.assert_array_equal(io.read_png(sio.optimize_dir / f"{k}.png"), ref)
    assert io.read_png(sio.optimize_dir / "tex.png").shape == (512, 512, 3)
    sio.save_images_th(imgs, sio.rerender_dir)
    saved = C.save_images(imgs.cpu().numpy())
    for k, i in enumerate(idx):
        np.testing.assert_array_equal(io.read_png(sio.rerender_dir / f"{i:02d}.png"), saved[k])
    # imread / imwrite with the reference's flags and layouts
    n = io.imread(ref_dir / "nom.png", "normal", (100, 80))
    assert n.shape == (80, 100, 3)
    np.testing.assert_array_equal(n.cpu().numpy(), C.imread_u8(maps["nom"], "normal", (100, 80)))
    r = io.imread(ref_dir / "rgh.png", "rough")
    assert r.shape == (300, 300)
    io.imwrite(n, tmp_path / "n.png", "normal")
    assert io.read_png(tmp_path / "n.png").shape == (80, 100, 3)


def test_codec_argument_validation(pkg):
    lib = pkg._lib.load()
    assert lib.svbrdf_resize_lanczos4_u8(None, 4, 4, 3, None, 8, 8, None) == -1
    t = torch.zeros(16, dtype=torch.uint8, device=DEV)
    p = pkg._lib.ptr(t)
    assert lib.svbrdf_resize_lanczos4_u8(p, 4, 4, 5, p, 8, 8, None) == -1
    assert lib.svbrdf_resize_lanczos4_u8(p, 0, 4, 3, p, 8, 8, None) == -2
    assert lib.svbrdf_images_from_u8(p, 1, 2, 2, 4, 1, p, None) == -1      # srgb needs 3 channels
    assert lib.svbrdf_images_from_u8(p, 1, 2, 2, 3, 9, p, None) == -1      # unknown flag
    assert lib.svbrdf_textures_from_u8(p, p, p, p, 2, 4, p, None) == -1
    assert lib.svbrdf_textures_to_u8(p, 0, p, p, p, p, None) == -2


@pytest.mark.parametrize("res", [1, 2, 5, 64, 256])
def test_upsample2x_matches_torch_interpolate(pkg, res):
    """Stage chaining in memory: svbrdf_upsample2x_bilinear against
    torch.nn.functional.interpolate(scale_factor=2, mode="bilinear",
    align_corners=False) on the GPU and, in float64, on the CPU.  The maps
    lie in [-1, 1]; the kernel rounds once per operation, so it stays within
    a few float32 ulps of 1 of both."""
    rng = np.random.default_rng(res)
    tex = torch.tensor(rng.uniform(-1, 1, (1, 9, res, res)).astype(np.float32), device=DEV)
    out = pkg.upsample_textures(tex)
    ref = torch.nn.functional.interpolate(tex, scale_factor=2, mode="bilinear", align_corners=False)
    ref64 = torch.nn.functional.interpolate(tex.cpu().double(), scale_factor=2, mode="bilinear", align_corners=False)
    torch.cuda.synchronize()
    assert out.shape == (1, 9, 2 * res, 2 * res)
    assert (out - ref).abs().max().item() <= 4e-7
    assert (out.cpu().double() - ref64).abs().max().item() <= 2.5e-7
    # the edge rows and columns repeat the source's (torch clamps the source coordinate at 0)
    assert torch.equal(out[0, :, 0, 0], tex[0, :, 0, 0])


def test_upsample2x_padded_planes_and_validation(pkg):
    """Source and destination at padded plane strides (the next stage's
    svbrdf_plane_stride layout): the result matches the dense call and the
    padding between planes is left as it was; bad strides and sizes raise."""
    L = pkg._lib
    res = 96
    rng = np.random.default_rng(7)
    tex = torch.tensor(rng.uniform(-1, 1, (1, 9, res, res)).astype(np.float32), device=DEV)
    dense = pkg.upsample_textures(tex)
    ps_src = L.query("svbrdf_plane_stride", res)
    ps_dst = L.query("svbrdf_plane_stride", 2 * res)
    src = torch.full((9 * ps_src,), float("nan"), device=DEV)
    src.view(9, ps_src)[:, :res * res].copy_(tex.reshape(9, -1))
    dst = torch.full((9 * ps_dst,), 7.0, device=DEV)
    stream = L.stream_handle(DEV)
    L.call("svbrdf_upsample2x_bilinear", L.ptr(src), res, ps_src, L.ptr(dst), ps_dst, 9, stream)
    torch.cuda.synchronize()
    planes = dst.view(9, ps_dst)
    assert torch.equal(planes[:, :4 * res * res], dense.reshape(9, -1))
    assert bool((planes[:, 4 * res * res:] == 7.0).all())
    with pytest.raises(ValueError):
        L.call("svbrdf_upsample2x_bilinear", L.ptr(src), res, res * res - 1, L.ptr(dst), 0, 9, stream)
    with pytest.raises(ValueError):
        L.call("svbrdf_upsample2x_bilinear", L.ptr(src), res, 0, L.ptr(dst), 4 * res * res - 1, 9, stream)
    with pytest.raises(ValueError):
        L.call("svbrdf_upsample2x_bilinear", L.ptr(src), 0, 0, L.ptr(dst), 0, 9, stream)
    with pytest.raises(ValueError):
        L.call("svbrdf_upsample2x_bilinear", L.ptr(src), res, 0, L.ptr(dst), 0, 65, stream)
    with pytest.raises(ValueError):
        pkg.upsample_textures(tex.cpu())


@pytest.mark.parametrize("res", [2, 4, 6, 96, 512])
def test_upsample2x_vector_path_equals_scalar_path(pkg, res):
    """Even resolutions with aligned destination rows take the strip kernel
    (4 source rows x 2 columns per thread; res 2 and 6 end in a partial
    strip); a destination one float off alignment takes the one output per
    thread kernel.  Both evaluate the same expressions, so they
    agree bit for bit, NaN and infinity propagation included (a NaN in the
    first column or row passes through the weight-0 neighbour term, as in
    torch)."""
    L = pkg._lib
    rng = np.random.default_rng(100 + res)
    src = torch.tensor(rng.uniform(-1, 1, (9, res, res)).astype(np.float32), device=DEV)
    src[0, 0, 1] = float("nan")
    src[1, 1, 0] = float("inf")
    src[2, res - 1, res - 1] = float("nan")
    n_out = 4 * res * res
    stream = L.stream_handle(DEV)
    vec = torch.full((9 * n_out,), 7.0, device=DEV)
    L.call("svbrdf_upsample2x_bilinear", L.ptr(src), res, 0, L.ptr(vec), 0, 9, stream)
    buf = torch.full((9 * n_out + 4,), 7.0, device=DEV)
    off = buf[1:1 + 9 * n_out]          # 4-byte offset: not 16-byte aligned
    L.call("svbrdf_upsample2x_bilinear", L.ptr(src), res, 0, L.ptr(off), 0, 9, stream)
    ref = torch.nn.functional.interpolate(src[None], scale_factor=2, mode="bilinear", align_corners=False)[0]
    torch.cuda.synchronize()
    a, b = vec.cpu().numpy(), off.cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert float(buf[0]) == 7.0 and float(buf[-3]) == 7.0
    r = ref.reshape(-1).cpu().numpy()
    fin = np.isfinite(r)
    assert np.array_equal(np.isnan(a), np.isnan(r)) and np.array_equal(np.isinf(a), np.isinf(r))
    assert np.abs(a[fin] - r[fin]).max() <= 4e-7


def test_upsample2x_refuses_overlapping_buffers(pkg):
    L = pkg._lib
    res = 8
    buf = torch.zeros(9 * 5 * res * res, device=DEV)
    stream = L.stream_handle(DEV)
    with pytest.raises(ValueError, match="overlap"):
        L.call("svbrdf_upsample2x_bilinear", L.ptr(buf), res, 0, L.ptr(buf), 0, 9, stream)
    with pytest.raises(ValueError, match="overlap"):
        # the destination starts inside the last source plane
        L.call("svbrdf_upsample2x_bilinear", L.ptr(buf), res, 0, L.ptr(buf[8 * res * res + 4:]), 0, 9, stream)
    # adjacent, not overlapping: accepted
    L.call("svbrdf_upsample2x_bilinear", L.ptr(buf), res, 0, L.ptr(buf[9 * res * res:]), 0, 9, stream)
    torch.cuda.synchronize()
