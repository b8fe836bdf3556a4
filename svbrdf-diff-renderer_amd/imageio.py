"""8-bit image I/O for maps and targets: src/imageio.py with the per-pixel
work on the GPU.

The reference reads and writes PNG files with cv2 (BGR order in memory) and
does the per-pixel work in numpy on the host.  Here the PNG container stays
on the host: files are read with Pillow and written by :func:`encode_png`
(Sub-filtered scanlines deflated in parallel pieces on a thread pool, zlib
level 1 as cv2's default); the arrays keep the file's RGB order, so the
reference's BGR<->RGB swaps cancel.  Everything per pixel runs in
libsvbrdf_hip.so:

* :func:`imresize` -- ``cv2.resize(..., INTER_LANCZOS4)`` of an 8-bit image
  (src/imageio.py:75-76), ``svbrdf_resize_lanczos4_u8``;
* :func:`imread` -- src/imageio.py:11-51: resize, ``u / 255`` and the
  "srgb" / "rough" / "normal" flags, ``svbrdf_images_from_u8``; returns a
  float32 CUDA tensor [H, W, C] (or [H, W] for "rough"), as the reference
  returns a numpy array of that shape;
* :func:`imwrite` -- src/imageio.py:54-72: clip per flag, ``* 255``,
  truncate, ``svbrdf_images_to_u8``;
* :func:`imconcat`, :func:`img9to1`, :func:`tex4to1` -- the montages of
  src/imageio.py:79-113.

Only 8-bit PNGs are handled (the reference's data are 8-bit; its uint16
branch, src/imageio.py:19-20, has no 16-bit RGB counterpart in Pillow).
"""
import concurrent.futures
import os
import pathlib
import struct
import zlib

import numpy as np
import torch
from PIL import Image

from . import _lib

FLAGS = {None: _lib.IMREAD_NONE, "srgb": _lib.IMREAD_SRGB, "rough": _lib.IMREAD_ROUGH,
         "normal": _lib.IMREAD_NORMAL}


def _flag(flag):
    try:
        return FLAGS[flag]
    except KeyError:
        raise ValueError(f"unknown flag {flag!r}; expected one of {list(FLAGS)}") from None


def read_png(filename):
    """The stored bytes of an 8-bit PNG: [H, W] (grey) or [H, W, C], RGB order."""
    with Image.open(str(filename)) as im:
        if im.mode in ("I", "I;16", "I;16B", "F"):
            raise ValueError(f"{filename}: only 8-bit images are supported, got mode {im.mode}")
        if im.mode == "P":
            im = im.convert("RGBA" if "transparency" in im.info else "RGB")
        if im.mode not in ("L", "LA", "RGB", "RGBA"):
            im = im.convert("RGB")
        return np.asarray(im, dtype=np.uint8).copy()


# zlib level 1: OpenCV's default speed setting for cv2.imwrite's PNGs
# (IMWRITE_PNG_COMPRESSION), which the reference writes with; the pixels are
# the same at any level.
PNG_COMPRESS_LEVEL = 1
_PNG_COLOR_TYPE = {1: 0, 2: 4, 3: 2, 4: 6}      # channels -> grey, grey+alpha, RGB, RGBA
_CHUNK_BYTES = 1 << 18                         # deflate work unit
_POOL = None


def _pool():
    global _POOL
    if _POOL is None:
        _POOL = concurrent.futures.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1))
    return _POOL


def _png_chunk(kind, data):
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def _png_pieces(arr):
    """Sub-filtered scanlines of [H, W] or [H, W, 1..4] uint8, cut into deflate
    pieces of whole rows: (IHDR payload, filtered bytes, pieces)."""
    arr = np.ascontiguousarray(arr, dtype=np.uint8)
    if arr.ndim == 2:
        arr = arr[:, :, None]
    if arr.ndim != 3 or arr.shape[2] not in _PNG_COLOR_TYPE or arr.shape[0] < 1 or arr.shape[1] < 1:
        raise ValueError(f"cannot write an 8-bit PNG of shape {arr.shape}")
    h, w, c = arr.shape
    rows = arr.reshape(h, w * c)
    raw = np.empty((h, w * c + 1), dtype=np.uint8)
    raw[:, 0] = 1                                   # filter type 1 (Sub)
    raw[:, 1:c + 1] = rows[:, :c]
    np.subtract(rows[:, c:], rows[:, :-c], out=raw[:, c + 1:])
    data = memoryview(raw.reshape(-1))
    step = max(1, _CHUNK_BYTES // raw.shape[1]) * raw.shape[1]
    pieces = [data[i:i + step] for i in range(0, len(data), step)]
    return struct.pack(">IIBBBBB", w, h, 8, _PNG_COLOR_TYPE[c], 0, 0, 0), data, pieces


def _deflate(piece, last):
    z = zlib.compressobj(PNG_COMPRESS_LEVEL, zlib.DEFLATED, -15)
    return z.compress(piece) + z.flush(zlib.Z_FINISH if last else zlib.Z_FULL_FLUSH)


def _png_parts(ihdr, data, body):
    """The file as a list of byte strings: the deflated pieces are not copied
    into one buffer, and the IDAT checksum runs over them in turn."""
    zhead = b"\x78\x01"
    adler = struct.pack(">I", zlib.adler32(data) & 0xFFFFFFFF)
    crc = zlib.crc32(zhead, zlib.crc32(b"IDAT"))
    for piece in body:
        crc = zlib.crc32(piece, crc)
    crc = zlib.crc32(adler, crc)
    size = len(zhead) + sum(len(piece) for piece in body) + len(adler)
    return [b"\x89PNG\r\n\x1a\n", _png_chunk(b"IHDR", ihdr), struct.pack(">I", size), b"IDAT", zhead, *body,
            adler, struct.pack(">I", crc & 0xFFFFFFFF), _png_chunk(b"IEND", b"")]


def _pmap(fn, items):
    items = list(items)
    return list(_pool().map(fn, items)) if len(items) > 1 else [fn(x) for x in items]


def _encode_parts(arrays):
    arrays = list(arrays)
    prepared = _pmap(_png_pieces, arrays)
    jobs = [(i, k) for i, (_, _, pieces) in enumerate(prepared) for k in range(len(pieces))]

    def run(job):
        i, k = job
        pieces = prepared[i][2]
        return _deflate(pieces[k], k == len(pieces) - 1)
    bodies = [[] for _ in arrays]
    for (i, _), out in zip(jobs, _pmap(run, jobs)):
        bodies[i].append(out)
    return _pmap(lambda i: _png_parts(prepared[i][0], prepared[i][1], bodies[i]), range(len(arrays)))


def encode_pngs(arrays):
    """PNG bytes of several [H, W] or [H, W, 1..4] uint8 images.  Every
    scanline uses the Sub filter (each byte minus the one a pixel to its left,
    mod 256: one numpy subtraction per image), and the filtered rows of all
    images are deflated in independent pieces on one thread pool (zlib and
    numpy release the GIL), each image's pieces joined as one zlib stream:
    full flushes between them, one header, and the Adler-32 of the whole
    stream -- the parallel-deflate layout any inflater reads."""
    return [b"".join(parts) for parts in _encode_parts(arrays)]


def encode_png(arr):
    """PNG bytes of one [H, W] or [H, W, 1..4] uint8 image (see encode_pngs)."""
    return encode_pngs([arr])[0]


def write_png(filename, arr):
    """Writes [H, W] or [H, W, 3 | 4] uint8 (RGB order) as an 8-bit PNG."""
    write_pngs([(filename, arr)])


def write_pngs(items):
    """Writes several (filename, array) PNGs, all encoded and written together
    on the thread pool.  Returns when all are on disk."""
    items = list(items)
    parts = _encode_parts(a for _, a in items)

    def write(i):
        with open(items[i][0], "wb") as f:
            f.writelines(parts[i])
    _pmap(write, range(len(items)))


def read_pngs(filenames):
    """read_png of several files on the thread pool (Pillow's decoder releases
    the GIL), in order."""
    filenames = list(filenames)
    if len(filenames) < 2:
        return [read_png(f) for f in filenames]
    return list(_pool().map(read_png, filenames))


def _device(device):
    dev = torch.device(device)
    if dev.type != "cuda":
        raise RuntimeError("svbrdf_amd image I/O runs on the GPU only; got device %r" % (str(dev),))
    return dev


def to_device_u8(arr, device="cuda"):
    """uint8 array or tensor -> contiguous uint8 CUDA tensor."""
    if isinstance(arr, torch.Tensor):
        t = arr
    else:
        t = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.uint8))
    return t.to(device=_device(device), dtype=torch.uint8).contiguous()


def _hwc(t):
    return t if t.dim() == 3 else t.unsqueeze(-1)


def imresize(im, dim):
    """cv2.resize(im, dim, interpolation=cv2.INTER_LANCZOS4) for an 8-bit
    [H, W] or [H, W, C] CUDA tensor; dim = (width, height) as cv2 takes it."""
    if im.dtype != torch.uint8 or im.device.type != "cuda":
        raise ValueError("imresize takes a uint8 CUDA tensor (8-bit images)")
    src = _hwc(im.contiguous())
    h, w, c = src.shape
    dw, dh = int(dim[0]), int(dim[1])
    out = torch.empty((dh, dw, c), dtype=torch.uint8, device=im.device)
    _lib.call("svbrdf_resize_lanczos4_u8", _lib.ptr(src), h, w, c, _lib.ptr(out), dh, dw,
              _lib.stream_handle(im.device))
    return out if im.dim() == 3 else out[:, :, 0]


def decode(u8, flag=None):
    """imread's arithmetic after the resize on [n, H, W, C] (or [H, W, C])
    uint8 CUDA images -> float32 [n, C', H, W] planar (C' = 1 for "rough")."""
    single = u8.dim() == 3
    x = u8.unsqueeze(0) if single else u8
    if x.dim() != 4:
        raise ValueError("decode takes [n, H, W, C] or [H, W, C] images")
    x = x.contiguous()
    n, h, w, c = x.shape
    f = _flag(flag)
    cout = 1 if f == _lib.IMREAD_ROUGH else (3 if f == _lib.IMREAD_NORMAL else c)
    out = torch.empty((n, cout, h, w), dtype=torch.float32, device=x.device)
    _lib.call("svbrdf_images_from_u8", _lib.ptr(x), n, h, w, c, f, _lib.ptr(out), _lib.stream_handle(x.device))
    return out[0] if single else out


def encode(images, flag=None):
    """imwrite's arithmetic before the PNG encode on [n, C, H, W] (or
    [C, H, W]) float32 CUDA images -> uint8 [n, H, W, C]."""
    single = images.dim() == 3
    x = images.unsqueeze(0) if single else images
    x = x.detach().to(torch.float32).contiguous()
    n, c, h, w = x.shape
    out = torch.empty((n, h, w, c), dtype=torch.uint8, device=x.device)
    _lib.call("svbrdf_images_to_u8", _lib.ptr(x), n, h, w, c, _flag(flag), _lib.ptr(out),
              _lib.stream_handle(x.device))
    return out[0] if single else out


def imread(filename, flag=None, dim=None, device="cuda"):
    """src/imageio.py:11-51.  Returns a float32 CUDA tensor [H, W, C] ([H, W]
    for "rough" and for grey images without a flag)."""
    u = read_png(filename)
    f = _flag(flag)
    if f in (_lib.IMREAD_SRGB, _lib.IMREAD_NORMAL) and (u.ndim != 3 or u.shape[2] != 3):
        raise ValueError(f"{filename} should be a 3 channel image")
    if f == _lib.IMREAD_ROUGH and not (u.ndim == 2 or (u.ndim == 3 and u.shape[2] == 3)):
        raise ValueError(f"{filename} should be a 3 or 1 channel image")
    t = to_device_u8(u, device)
    if dim is not None:
        t = imresize(t, dim)
    out = decode(_hwc(t), flag)           # [C', H, W]
    if f == _lib.IMREAD_ROUGH or (u.ndim == 2 and out.shape[0] == 1):
        return out[0]
    return out.permute(1, 2, 0)


def imwrite(im, filename, flag=None, dim=None):
    """src/imageio.py:54-72 for a float [H, W, C] or [H, W] image (CUDA tensor
    or array).  The reference's optional float resize before writing is not
    used on its hot path and is not provided."""
    if dim is not None:
        raise NotImplementedError("imwrite(dim=...) (a float-image resize) is not provided")
    t = torch.as_tensor(im)
    if t.device.type != "cuda":
        t = t.to(_device("cuda"))
    chw = t.unsqueeze(0) if t.dim() == 2 else t.permute(2, 0, 1)
    u = encode(chw, flag)
    write_png(filename, u[:, :, 0].cpu().numpy() if t.dim() == 2 else u.cpu().numpy())


def imconcat(im_list, size=(2, 2)):
    """src/imageio.py:79-90: rows of `size[0]` images, `size[1]` rows."""
    w, h = size
    rows = [np.concatenate(im_list[i * w:(i + 1) * w], axis=1) for i in range(h)]
    return np.concatenate(rows, axis=0)


def montage9(ims):
    """The 3 x 3 montage of img9to1 from the 9 stored images (bytes)."""
    return imconcat(list(ims), (3, 3))


def montage_tex(nom, dif, spe, rgh):
    """The 2 x 2 montage of tex4to1 from the stored maps (bytes); a grey
    roughness map is made 3-channel."""
    if rgh.ndim == 2:
        rgh = np.repeat(rgh[:, :, None], 3, axis=2)
    return imconcat([nom, dif, spe, rgh])


def img9to1(folder):
    """src/imageio.py:93-99: the 9 views 00..08.png as a 3 x 3 montage all.png.
    The reference passes every pixel through imread's u / 255 and imwrite's
    (v * 255).astype(uint8); in float32 that returns each of the 256 levels
    unchanged, so the montage holds the stored bytes."""
    folder = pathlib.Path(folder)
    write_png(folder / "all.png", montage9(read_png(folder / f"{i:02d}.png") for i in range(9)))


def tex4to1(folder):
    """src/imageio.py:102-113: normal, diffuse, specular, roughness (grey made
    3-channel) as a 2 x 2 montage tex.png (stored bytes, as in img9to1)."""
    folder = pathlib.Path(folder)
    maps = [read_png(folder / f"{n}.png") for n in ("nom", "dif", "spe", "rgh")]
    write_png(folder / "tex.png", montage_tex(*maps))
