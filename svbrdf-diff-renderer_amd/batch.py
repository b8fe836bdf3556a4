"""Many materials, one launch per iteration.

The reference optimises one material per process (src/scripts.py:67-97).
``BatchedStep`` runs the same per-pixel iteration (clamp, render, loss,
adjoint, Adam; src/svbrdf.py:58-71) for a whole batch of materials of
uneven resolution and view count in one kernel launch over a table of
pixel tiles in cost order (views x pixels, costliest first).  Two launch
forms: one block per 256-pixel tile (blocks_per_cu = -1, the hardware
dispatcher balances the load), or a persistent grid whose blocks draw
tiles of 1,024-16,384 pixels from an atomic ticket counter.

The launch shape -- form, tile size, pixels per thread, views per loop trip
and blocks per compute unit -- is chosen by :meth:`BatchedStep.autotune`,
which times each candidate on the batch itself (state saved and restored
around the trials) and caches the winner per (device, loss, batch shape) in
a JSON file.
"""
import ctypes
import hashlib
import json
import os
import pathlib

import numpy as np
import torch

from . import _lib
from .optim import _STATE_DTYPE, _planes, loss_mode

DEFAULT_CONFIG = {"tile_pixels": 256, "pixels_per_thread": 1, "views_per_trip": 1, "blocks_per_cu": -1}


class Material(ctypes.Structure):
    """ctypes view of ``svbrdf_material`` (include/svbrdf_hip.h)."""
    _fields_ = [("textures", ctypes.c_void_p), ("exp_avg", ctypes.c_void_p), ("exp_avg_sq", ctypes.c_void_p),
                ("targets_prepared", ctypes.c_void_p), ("camera_pos", ctypes.c_void_p),
                ("light_pos", ctypes.c_void_p), ("light_pow", ctypes.c_void_p), ("res", ctypes.c_int),
                ("n_views", ctypes.c_int), ("size", ctypes.c_float), ("tile_begin", ctypes.c_uint),
                ("n_tiles", ctypes.c_uint), ("plane_stride", ctypes.c_uint), ("texture_stride", ctypes.c_uint),
                ("loss_scale", ctypes.c_float), ("light_adam", ctypes.c_void_p)]


def _cache_path():
    root = os.environ.get("SVBRDF_AMD_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "svbrdf_amd")
    return pathlib.Path(root) / "autotune.json"


def _compute_units(device):
    return torch.cuda.get_device_properties(device).multi_processor_count


class BatchedStep:
    """A batch of independent per-pixel optimisation problems.

    problems: sequence of (renderer, textures, targets) with renderer a
    Microfacet (geometry, light power, size), textures a contiguous fp32
    CUDA tensor of 9*R*R values (updated in place, or with own_textures=True
    copied into a buffer at the padded plane stride that write_back()
    returns) and targets [V,3,R,R].  All materials share lr, betas, eps and
    the loss, and step together.  With optim_light=True each material's light
    power is optimised too (SvbrdfOptim.optim(optim_light=True) per material,
    src/svbrdf.py:48-50,58-59): the renderer's own light_pow tensor (a
    contiguous float32 [3] CUDA tensor) is updated in place, as FusedStep
    does, with its Adam moments in ``light_adam[m]``.
    """

    def __init__(self, problems, lr, loss="log_l2", betas=(0.9, 0.999), eps=1e-8, history=0, config=None,
                 own_textures=False, optim_light=False):
        if not problems:
            raise ValueError("empty batch")
        self.device = _lib.cuda_device(problems[0][0].device, "BatchedStep")
        self.lr, self.beta1, self.beta2, self.eps = float(lr), float(betas[0]), float(betas[1]), float(eps)
        self.mode = loss_mode(loss)
        s = _lib.stream_handle(self.device)
        self.renderers, self.textures, self.exp_avg, self.exp_avg_sq, self.targets = [], [], [], [], []
        self.plane_strides, self.sources = [], []
        for r, tex, tgt in problems:
            if tex.device != self.device or tex.dtype != torch.float32 or not tex.is_contiguous():
                raise ValueError("textures must be contiguous float32 tensors on one CUDA device")
            if tex.numel() != 9 * r.res * r.res or tuple(tgt.shape) != (r.n_of_imgs, 3, r.res, r.res):
                raise ValueError(f"material of res {r.res} / {r.n_of_imgs} views: shape mismatch")
            tg = tgt.detach().to(self.device, torch.float32).contiguous()
            dense = torch.empty_like(tg)
            _lib.call("svbrdf_prepare_targets", _lib.ptr(tg), tg.numel(), self.mode, _lib.ptr(dense), s)
            # moments and prepared targets at the library's padded plane stride
            ps = _lib.query("svbrdf_plane_stride", r.res)
            moments = torch.zeros((2, 9 * ps), dtype=torch.float32, device=self.device)
            prepared = _planes(torch.zeros(3 * r.n_of_imgs * ps, dtype=torch.float32, device=self.device),
                               tg.shape, ps)
            prepared.copy_(dense)
            self.renderers.append(r)
            self.sources.append(tex)
            if own_textures:
                own = _planes(torch.zeros(9 * ps, dtype=torch.float32, device=self.device), tex.shape, ps)
                with torch.no_grad():
                    own.copy_(tex)
                tex = own
            self.textures.append(tex)
            self.plane_strides.append(ps)
            self.exp_avg.append(_planes(moments[0], tex.shape, ps))
            self.exp_avg_sq.append(_planes(moments[1], tex.shape, ps))
            self.targets.append(prepared)
        self.optim_light = bool(optim_light)
        if self.optim_light:
            for r in self.renderers:
                lp = r.light_pow
                if (lp.device != self.device or lp.dtype != torch.float32 or lp.numel() != 3
                        or not lp.is_contiguous()):
                    raise ValueError("optim_light needs each renderer's light_pow as a contiguous float32 [3] "
                                     "tensor on the batch's device")
            self.light_pow = [r.light_pow for r in self.renderers]
        else:
            self.light_pow = [r.light_pow.detach().to(self.device, torch.float32).reshape(3).contiguous()
                              for r in self.renderers]
        # Adam's moments of each material's light power: exp_avg[3], exp_avg_sq[3]
        self.light_adam = torch.zeros((len(problems), 8), dtype=torch.float32, device=self.device)
        self.n = len(problems)
        self.own_textures = bool(own_textures)
        nstate = _lib.query("svbrdf_step_state_bytes")
        self.state = torch.zeros((nstate + 15) // 16 * 4, dtype=torch.float32, device=self.device)
        _lib.call("svbrdf_step_state_init", _lib.ptr(self.state), self.beta1, self.beta2, s)
        self.losses_dev = torch.zeros(self.n, dtype=torch.float32, device=self.device)
        self.history = torch.zeros((max(int(history), 0), self.n), dtype=torch.float32, device=self.device)
        self.steps_launched = 0
        self.graph = None
        self.graph_steps = 0
        self.cu = _compute_units(self.device)
        self.configure(dict(DEFAULT_CONFIG, **(config or {})))

    # ------------------------------------------------------------------ plan
    def signature(self):
        shape = sorted((r.res, r.n_of_imgs) for r in self.renderers)
        h = hashlib.sha1(json.dumps(shape).encode()).hexdigest()[:16]
        light = "|light" if self.optim_light else ""
        return f"{torch.cuda.get_device_name(self.device)}|loss{self.mode}{light}|{h}"

    def configure(self, cfg):
        """Plans tiles for ``cfg`` and builds the device tables."""
        cfg = dict(cfg)
        npix_t, unroll = int(cfg["pixels_per_thread"]), int(cfg["views_per_trip"])
        tile = int(cfg["tile_pixels"])
        if tile % (256 * npix_t):
            raise ValueError(f"tile_pixels {tile} is not a multiple of {256 * npix_t}")
        bpc = int(cfg.get("blocks_per_cu") or 0)
        if bpc < 0:
            # one block per 256-pixel tile, no persistent loop
            if tile != 256 or npix_t != 1:
                raise ValueError("blocks_per_cu=-1 (one block per tile) needs tile_pixels=256, pixels_per_thread=1")
            cfg["blocks_per_cu"] = -1
        else:
            max_bpc = _lib.query("svbrdf_batch_max_blocks_per_cu", self.mode, int(self.optim_light), npix_t, unroll)
            if max_bpc < 1:
                raise RuntimeError("svbrdf_batch_max_blocks_per_cu failed: "
                                   + _lib.load().svbrdf_last_error().decode())
            bpc = max_bpc if bpc == 0 else min(bpc, max_bpc)
            cfg["blocks_per_cu"] = bpc
        table = (Material * self.n)()
        for i, r in enumerate(self.renderers):
            table[i] = Material(self.textures[i].data_ptr(), self.exp_avg[i].data_ptr(),
                                self.exp_avg_sq[i].data_ptr(), self.targets[i].data_ptr(),
                                r.camera_pos.data_ptr(), r.light_pos.data_ptr(), self.light_pow[i].data_ptr(),
                                r.res, r.n_of_imgs, r.size, 0, 0, self.plane_strides[i],
                                self.plane_strides[i] if self.own_textures else 0, 0.0,
                                self.light_adam[i].data_ptr() if self.optim_light else None)
        n_tiles = _lib.query("svbrdf_batch_tile_count", table, self.n, tile)
        if n_tiles == 0:
            raise ValueError("invalid batch for tile planning")
        order = (ctypes.c_uint * n_tiles)()
        tile_mat = (ctypes.c_uint * n_tiles)()
        _lib.call("svbrdf_batch_plan", table, self.n, tile, order, tile_mat, n_tiles)
        self.table_host = table
        self.table = torch.frombuffer(bytearray(bytes(table)), dtype=torch.uint8).to(self.device)
        self.order = torch.frombuffer(bytearray(bytes(order)), dtype=torch.int32).to(self.device)
        self.tile_mat = torch.frombuffer(bytearray(bytes(tile_mat)), dtype=torch.int32).to(self.device)
        nws = _lib.query("svbrdf_batch_workspace_bytes", n_tiles)
        self.workspace = torch.zeros(nws // 4 + 64, dtype=torch.float32, device=self.device)
        off = (-self.workspace.data_ptr()) % 256
        self.ws_ptr = self.workspace.data_ptr() + off
        self.ws_bytes = self.workspace.numel() * 4 - off
        self.n_tiles = n_tiles
        self.config = cfg
        self.grid = 0 if bpc < 0 else bpc * self.cu
        self.graph = None
        return cfg

    # ----------------------------------------------------------------- steps
    def _launch(self, stream):
        c = self.config
        hist = self.history if self.history.numel() else None
        _lib.call("svbrdf_batch_step", _lib.ptr(self.table), self.n, _lib.ptr(self.order), _lib.ptr(self.tile_mat),
                  self.n_tiles, c["tile_pixels"], c["pixels_per_thread"], c["views_per_trip"], self.grid,
                  self.mode, int(self.optim_light), self.lr, self.beta1, self.beta2, self.eps, _lib.ptr(self.state),
                  _lib.ptr(self.losses_dev), _lib.ptr(hist), self.history.shape[0], ctypes.c_void_p(self.ws_ptr),
                  self.ws_bytes, stream)

    def step(self):
        self._launch(_lib.stream_handle(self.device))
        self.steps_launched += 1

    def capture(self, steps=1):
        """Records `steps` consecutive iterations in one HIP graph; every
        replay() runs that many (a graph launch costs tens of microseconds)."""
        if steps < 1:
            raise ValueError("steps must be >= 1")
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            g.capture_begin()
            for _ in range(steps):
                self._launch(ctypes.c_void_p(s.cuda_stream))
            g.capture_end()
        torch.cuda.current_stream().wait_stream(s)
        self.graph = g
        self.graph_steps = steps
        return g

    def replay(self):
        self.graph.replay()
        self.steps_launched += self.graph_steps

    def write_back(self):
        """Copies each material's maps into the tensor it was given when the
        batch keeps its own copies (own_textures=True); asynchronous."""
        if self.own_textures:
            with torch.no_grad():
                for src, own in zip(self.sources, self.textures):
                    src.copy_(own)

    def losses(self):
        """Per-material losses of the last step (synchronises)."""
        return self.losses_dev.cpu().numpy()

    def loss_history(self):
        n = min(self.steps_launched, self.history.shape[0])
        return self.history[:n].cpu().numpy()

    def read_state(self):
        raw = self.state.detach().cpu().numpy().view(np.uint8)[:_STATE_DTYPE.itemsize]
        return raw.view(_STATE_DTYPE)[0]

    # -------------------------------------------------------------- autotune
    def _snapshot(self):
        return ([t.clone() for t in self.textures], [m.clone() for m in self.exp_avg],
                [v.clone() for v in self.exp_avg_sq], self.state.clone(), self.losses_dev.clone(),
                self.history.clone(), self.steps_launched, [p.clone() for p in self.light_pow],
                self.light_adam.clone())

    def _restore(self, snap):
        with torch.no_grad():
            for dst, src in zip(self.textures, snap[0]):
                dst.copy_(src)
            for dst, src in zip(self.exp_avg, snap[1]):
                dst.copy_(src)
            for dst, src in zip(self.exp_avg_sq, snap[2]):
                dst.copy_(src)
            self.state.copy_(snap[3])
            self.losses_dev.copy_(snap[4])
            self.history.copy_(snap[5])
            for dst, src in zip(self.light_pow, snap[7]):
                dst.copy_(src)
            self.light_adam.copy_(snap[8])
        self.steps_launched = snap[6]

    @staticmethod
    def candidates():
        out = []
        for unroll in (1, 3):
            # one block per 256-pixel tile (blocks_per_cu = -1)
            out.append({"tile_pixels": 256, "pixels_per_thread": 1, "views_per_trip": unroll, "blocks_per_cu": -1})
        for npix_t in (1, 2):
            for unroll in (1, 3):
                for tile in (1024, 4096, 16384):
                    for bpc in (0, 2):
                        out.append({"tile_pixels": tile, "pixels_per_thread": npix_t, "views_per_trip": unroll,
                                    "blocks_per_cu": bpc})
        return out

    def autotune(self, candidates=None, iters=5, use_cache=True, verbose=False, rounds=2):
        """Times each candidate launch shape on this batch and keeps the
        fastest (best of `rounds` passes over the candidates, `iters` steps
        each); the optimisation state is restored afterwards.  Returns
        (config, {config key: ms per step})."""
        key = self.signature()
        path = _cache_path()
        if use_cache and candidates is None:
            try:
                cached = json.loads(path.read_text()).get(key)
            except (OSError, ValueError):
                cached = None
            if cached:
                return self.configure(cached), {}
        snap = self._snapshot()
        shapes = {}
        for cand in candidates or self.candidates():
            cfg = self.configure(cand)
            shapes.setdefault(json.dumps(cfg, sort_keys=True), cfg)
        timings = {}
        for _ in range(max(1, rounds)):
            for k, cfg in shapes.items():
                # timed as the batch runs: replays of a graph of `iters` steps
                self.configure(cfg)
                g = self.capture(iters)
                g.replay()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                torch.cuda.synchronize(self.device)
                e0.record()
                g.replay()
                e1.record()
                torch.cuda.synchronize(self.device)
                ms = e0.elapsed_time(e1) / iters
                del g
                self.graph = None
                timings[k] = min(ms, timings.get(k, float("inf")))
                if verbose:
                    print(f"autotune {k}: {ms:.3f} ms", flush=True)
                self._restore(snap)
        best = (timings[min(timings, key=timings.get)], shapes[min(timings, key=timings.get)])
        self._restore(snap)
        cfg = self.configure(best[1])
        if use_cache:
            try:
                path.parent.mkdir(parents=True, exist_ok=True)
                data = json.loads(path.read_text()) if path.exists() else {}
                data[key] = cfg
                path.write_text(json.dumps(data, indent=1, sort_keys=True))
            except (OSError, ValueError):
                pass
        return cfg, timings
