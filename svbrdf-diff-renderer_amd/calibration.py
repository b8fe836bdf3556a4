"""Flash calibration in the reference's JSON schema.

Mirrors SvbrdfIO.__init__ / load_calibration_th (src/svbrdf.py:87-148): the
keys ``camera_pos``, ``light_pos`` ([N,3], cm), ``light_pow`` ([3]),
``idx`` (the views to use, in order) and ``im_size`` (plane width, cm) give
``cl = [camera_pos[idx], light_pos[idx], light_pow]`` as float32 tensors, the
third argument of Microfacet(res, n, size, cl, device).
"""
import json
import pathlib

import numpy as np
import torch


def load_calibration(json_path, device="cpu"):
    """Returns (cl, im_size, idx) from a calibration JSON (src/svbrdf.py:97-148)."""
    path = pathlib.Path(json_path)
    if not path.exists():
        raise FileNotFoundError(f"{path} does not exist")
    data = json.loads(path.read_text())
    missing = [k for k in ("camera_pos", "light_pos", "light_pow", "idx") if k not in data]
    if missing:
        raise KeyError(f"{path}: missing {missing}")
    idx = [int(i) for i in data["idx"]]
    cam = np.asarray(data["camera_pos"], np.float32)[idx, :]
    light = np.asarray(data["light_pos"], np.float32)[idx, :]
    pow_ = np.asarray(data["light_pow"], np.float32)
    if cam.ndim != 2 or cam.shape[1] != 3 or light.shape != cam.shape or pow_.shape != (3,):
        raise ValueError(f"{path}: camera_pos/light_pos must be [N,3] and light_pow [3]")
    cl = [torch.from_numpy(cam).to(device), torch.from_numpy(light).to(device), torch.from_numpy(pow_).to(device)]
    return cl, float(data.get("im_size", 0.0)), idx
