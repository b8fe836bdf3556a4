"""SvbrdfIO: the reference's map / target / calibration I/O object
(src/svbrdf.py:86-221) with the per-pixel codec on the GPU.

Same constructor (a JSON path and a device), attributes (reference_dir,
target_dir, optimize_dir, rerender_dir, im_size, idx, n_of_imgs, camera_pos,
light_pos, light_pow, cl) and methods.  Where the reference prints an error
and calls exit() this raises (FileNotFoundError / ValueError).

* load_textures_th -- 4 PNG maps -> [1, 9, res, res]: 8-bit Lanczos-4 resize
  and the imread flags in ``svbrdf_resize_lanczos4_u8`` and
  ``svbrdf_textures_from_u8``;
* save_textures_th -- [1, 9, R, R] -> 4 PNG maps + tex.png:
  ``svbrdf_textures_to_u8`` (reconstruct_normal, clip, * 255, truncate);
* load_images_th / save_images_th -- the target / re-render images through
  ``svbrdf_images_from_u8`` / ``svbrdf_images_to_u8`` ("srgb");
* :func:`reload_textures` -- what save_textures_th followed by
  load_textures_th(res) gives, without files: the stage transition of the
  reference's coarse-to-fine runs (src/scripts.py:83-85 chains stages
  through the saved maps).
"""
import json
import pathlib

import numpy as np
import torch

from . import _lib
from .imageio import (decode, encode, img9to1, imresize, montage9, montage_tex, read_pngs, to_device_u8,
                      write_pngs)

MAP_FILES = {"normal": "nom.png", "diffuse": "dif.png", "specular": "spe.png", "roughness": "rgh.png"}


def _check_device(device):
    return _lib.cuda_device(device, "SvbrdfIO")


def calibration_arrays(camera_pos, light_pos, light_pow, idx):
    """The host half of load_calibration_th (src/svbrdf.py:134-148): the
    views listed in ``idx``, in that order, of camera_pos / light_pos ([N, 3],
    cm) and the per-RGB light_pow [3], as contiguous float32 arrays."""
    cam = np.asarray(camera_pos, np.float32)
    light = np.asarray(light_pos, np.float32)
    pow_ = np.asarray(light_pow, np.float32)
    if cam.ndim != 2 or cam.shape[1] != 3 or light.shape != cam.shape or pow_.shape != (3,):
        raise ValueError("camera_pos / light_pos must be [N, 3] and light_pow [3]")
    sel = [int(i) for i in idx]
    return np.ascontiguousarray(cam[sel, :]), np.ascontiguousarray(light[sel, :]), pow_


def _sized(u8, res):
    """imresize to res x res unless already that size (a copy in cv2)."""
    if u8.shape[0] == res and u8.shape[1] == res:
        return u8
    return imresize(u8, (res, res))


def textures_from_maps_u8(dif, nom, spe, rgh, res):
    """The four 8-bit maps (CUDA uint8, [H, W, 3]; roughness [H, W] or
    [H, W, 3]) -> textures [1, 9, res, res] as load_textures_th produces them."""
    dev = dif.device
    dif, nom, spe = (_sized(m, res).contiguous() for m in (dif, nom, spe))
    rgh = _sized(rgh, res).contiguous()
    for name, m in (("diffuse", dif), ("normal", nom), ("specular", spe)):
        if m.dim() != 3 or m.shape[2] != 3:
            raise ValueError(f"{name} map should be a 3 channel image")
    if not (rgh.dim() == 2 or (rgh.dim() == 3 and rgh.shape[2] == 3)):
        raise ValueError("roughness map should be a 3 or 1 channel image")
    rc = 1 if rgh.dim() == 2 else 3
    tex = torch.empty((1, 9, res, res), dtype=torch.float32, device=dev)
    _lib.call("svbrdf_textures_from_u8", _lib.ptr(dif), _lib.ptr(nom), _lib.ptr(spe), _lib.ptr(rgh), rc, res,
              _lib.ptr(tex), _lib.stream_handle(dev))
    return tex


def maps_u8_from_textures(textures):
    """textures [1, 9, R, R] or [9, R, R] -> (dif, nom, spe [R, R, 3], rgh
    [R, R]) uint8 CUDA tensors, as save_textures_th writes them."""
    t = textures.detach().to(torch.float32).reshape(9, textures.shape[-2], textures.shape[-1]).contiguous()
    res = t.shape[-1]
    if t.shape[-2] != res:
        raise ValueError("textures must be square")
    dev = t.device
    dif, nom, spe = (torch.empty((res, res, 3), dtype=torch.uint8, device=dev) for _ in range(3))
    rgh = torch.empty((res, res), dtype=torch.uint8, device=dev)
    _lib.call("svbrdf_textures_to_u8", _lib.ptr(t), res, _lib.ptr(dif), _lib.ptr(nom), _lib.ptr(spe), _lib.ptr(rgh),
              _lib.stream_handle(dev))
    return dif, nom, spe, rgh


def reload_textures(textures, res):
    """save_textures_th then load_textures_th(res), in device memory: 8-bit
    quantisation of the maps, Lanczos-4 resize to res, decode."""
    dif, nom, spe, rgh = maps_u8_from_textures(textures)
    return textures_from_maps_u8(dif, nom, spe, rgh, res)


class SvbrdfIO:
    """src/svbrdf.py:86-221."""

    def __init__(self, json_dir, device):
        self.device = _check_device(device)
        json_dir = pathlib.Path(json_dir)
        if not json_dir.exists():
            raise FileNotFoundError(f"{json_dir} does not exist")
        data = json.loads(json_dir.read_text())
        for key in ("reference_dir", "target_dir", "optimize_dir", "rerender_dir"):
            if key in data:
                setattr(self, key, json_dir.parent / data[key])
        if "im_size" in data:
            self.im_size = data["im_size"]
        if "idx" in data:
            self.idx = data["idx"]
            self.n_of_imgs = len(self.idx)
        if "camera_pos" in data:
            self.camera_pos = data["camera_pos"]
        if "light_pos" in data:
            self.light_pos = data["light_pos"]
        if "light_pow" in data:
            self.light_pow = data["light_pow"]
            self.load_calibration_th()

    def np_to_th(self, arr):
        return torch.from_numpy(arr).to(self.device)

    def th_to_np(self, arr):
        return arr.detach().cpu().numpy()

    def reconstruct_normal(self, texture):
        """src/svbrdf.py:126-132 on [N, 2, R, R] (torch ops on the device; the
        save path runs the same arithmetic inside svbrdf_textures_to_u8)."""
        nx = texture[:, 0].clamp(-1, 1)
        ny = texture[:, 1].clamp(-1, 1)
        nxy = (nx ** 2 + ny ** 2).clamp(0, 1)
        nz = (1 - nxy).sqrt()
        n = torch.stack((nx, ny, nz), 1)
        return n / n.norm(2.0, 1, keepdim=True)

    def load_calibration_th(self):
        """src/svbrdf.py:134-148."""
        cam, light, pow_ = calibration_arrays(self.camera_pos, self.light_pos, self.light_pow, self.idx)
        self.cl = [self.np_to_th(cam), self.np_to_th(light), self.np_to_th(pow_)]

    def load_textures_th(self, textures_dir, res):
        """src/svbrdf.py:150-166 -> [1, 9, res, res] in [-1, 1]."""
        d = pathlib.Path(textures_dir)
        if not d.exists():
            raise FileNotFoundError(f"{d} does not exist")
        names = list(MAP_FILES)
        arrays = read_pngs(d / MAP_FILES[k] for k in names)
        maps = {k: to_device_u8(a, self.device) for k, a in zip(names, arrays)}
        return textures_from_maps_u8(maps["diffuse"], maps["normal"], maps["specular"], maps["roughness"], res)

    def save_textures_th(self, textures_th, textures_dir):
        """src/svbrdf.py:168-189: nom / dif / spe / rgh.png and tex.png."""
        d = pathlib.Path(textures_dir)
        d.mkdir(parents=True, exist_ok=True)
        dif, nom, spe, rgh = (m.cpu().numpy() for m in maps_u8_from_textures(textures_th))
        # the montage of tex4to1 from the same bytes the map files hold; all
        # five files encoded at once
        write_pngs([(d / "nom.png", nom), (d / "dif.png", dif), (d / "spe.png", spe), (d / "rgh.png", rgh),
                    (d / "tex.png", montage_tex(nom, dif, spe, rgh))])

    def load_images_th(self, images_dir, res=256):
        """src/svbrdf.py:191-204 -> [n_of_imgs, 3, res, res] in [0, 1]."""
        d = pathlib.Path(images_dir)
        if not d.exists():
            raise FileNotFoundError(f"{d} does not exist")
        ims = []
        for i, u in zip(self.idx, read_pngs(d / f"{i:02d}.png" for i in self.idx)):
            if u.ndim != 3 or u.shape[2] != 3:
                raise ValueError(f"{d / f'{i:02d}.png'} should be a 3 channel image")
            ims.append(_sized(to_device_u8(u, self.device), res))
        return decode(torch.stack(ims), "srgb")

    def save_images_th(self, images_th, images_dir):
        """src/svbrdf.py:206-221: <idx>.png per view (+ all.png for 9 views)."""
        if images_th.shape[0] != self.n_of_imgs:
            raise ValueError(f"expected {self.n_of_imgs} images, got {images_th.shape[0]}")
        d = pathlib.Path(images_dir)
        d.mkdir(parents=True, exist_ok=True)
        u = encode(images_th, "srgb").cpu().numpy()
        items = [(d / f"{i:02d}.png", u[k]) for k, i in enumerate(self.idx)]
        if self.n_of_imgs == 9 and list(self.idx) == list(range(9)):
            # img9to1's montage of 00..08.png, from the bytes just encoded
            items.append((d / "all.png", montage9(u)))
        write_pngs(items)
        if self.n_of_imgs == 9 and list(self.idx) != list(range(9)):
            img9to1(d)
