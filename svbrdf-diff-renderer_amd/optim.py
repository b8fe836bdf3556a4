"""Per-pixel SVBRDF optimisation on the fused gfx950 step kernel.

``SvbrdfOptim`` mirrors the reference optimiser (src/svbrdf.py:15-83 on top
of src/optimization.py:10-42): same constructor, initialisers, target
loading and ``optim(epochs, lr, svbrdf_obj, optim_light)`` loop, with the
loop body -- clamp, render, loss, backward, Adam -- executed as ONE kernel
launch per iteration (:class:`FusedStep`).

``FusedStep`` owns the device-resident pieces of one optimisation problem:
the prepared targets, Adam's moment buffers, the step state and the
workspace (per-block partial sums of the last step and arrival counters).
Its ``step()`` is graph-capturable (no host scalars change between
iterations); ``capture(steps)`` records that many iterations in one HIP
graph and ``replay()`` runs them.
"""
import ctypes
import datetime
import pathlib

import numpy as np
import torch

from . import _lib

LOSSES = {"l2": _lib.LOSS_L2, "log_l2": _lib.LOSS_LOG_L2, "logl2": _lib.LOSS_LOG_L2}

# layout of svbrdf_step_state (include/svbrdf_hip.h)
_STATE_DTYPE = np.dtype([("beta1_pow", "<f8"), ("beta2_pow", "<f8"),
                         ("inv_bias_correction1", "<f8"), ("inv_bias_correction2_sqrt", "<f4"),
                         ("step", "<f4"), ("loss", "<f4"), ("pending_step", "<u4"),
                         ("reserved", "<u4", 2),
                         ("light_exp_avg", "<f4", 4), ("light_exp_avg_sq", "<f4", 4),
                         ("light_grad", "<f4", 4)])


def loss_mode(loss):
    if isinstance(loss, int):
        return loss
    try:
        return LOSSES[loss.lower()]
    except KeyError:
        raise ValueError(f"unknown loss {loss!r}; expected one of {sorted(LOSSES)}") from None


def _planes(buf, shape, stride):
    """[.., P, R, R] view of a flat buffer whose planes start `stride` floats
    apart (leading dimensions are plane groups, contiguous in P planes)."""
    shape = tuple(int(x) for x in shape)
    strides = [shape[-1], 1]
    acc = stride
    for d in reversed(shape[:-2]):
        strides.insert(0, acc)
        acc *= d
    return buf.as_strided(shape, strides)


def upsample_textures(textures):
    """The maps of the next coarse-to-fine stage: ``[1, 9, R, R]`` float32
    on the GPU -> ``[1, 9, 2R, 2R]``, with the semantics of
    ``torch.nn.functional.interpolate(textures, scale_factor=2,
    mode="bilinear", align_corners=False)``, computed by
    svbrdf_upsample2x_bilinear.  The reference chains stages through 8-bit
    PNG files and a Lanczos-4 resize instead (src/scripts.py:83-85; that
    route is :func:`svbrdf_io.reload_textures`)."""
    if textures.dim() != 4 or textures.shape[0] != 1 or textures.shape[2] != textures.shape[3]:
        raise ValueError(f"textures must be [1, C, R, R], got {tuple(textures.shape)}")
    if textures.dtype != torch.float32 or textures.device.type != "cuda":
        raise ValueError("textures must be float32 on a GPU (there is no CPU path)")
    src = textures.detach().contiguous()
    n_planes, res = src.shape[1], src.shape[2]
    out = torch.empty((1, n_planes, 2 * res, 2 * res), dtype=torch.float32, device=src.device)
    _lib.call("svbrdf_upsample2x_bilinear", _lib.ptr(src), res, 0, _lib.ptr(out), 0, n_planes,
              _lib.stream_handle(src.device))
    return out


def epoch_segments(epochs, save_every=100, sync_every=0, saving=True):
    """Splits an optimisation of `epochs` iterations at the epochs after which
    the host acts: [(first, last, save)], 0-based and inclusive.  `save` marks
    the reference's save points -- epoch 1, every `save_every`-th and the last
    (src/svbrdf.py:74) -- when `saving`; a segment also ends every
    `sync_every` epochs when that is set.  The iterations of a segment need no
    host work between them."""
    if save_every < 1:
        raise ValueError("save_every must be >= 1")
    out, first = [], 0
    for e in range(epochs):
        save = saving and ((e + 1) % save_every == 0 or e == 0 or e == epochs - 1)
        if save or (sync_every and (e + 1) % sync_every == 0) or e == epochs - 1:
            out.append((first, e, save))
            first = e + 1
    return out


class FusedStep:
    """One optimisation problem: textures (updated in place) against targets.

    textures: [1,9,R,R] or [9,R,R] float32 CUDA tensor, updated in place
              (with own_textures=True the step works on a padded copy,
              ``self.textures``, and write_back() updates this tensor).
    targets:  [V,3,R,R] float32 images in gamma space (what Microfacet.eval
              produces and what src/svbrdf.py:191-204 loads).
    renderer: a :class:`Microfacet` (geometry, light power, size).
    """

    def __init__(self, renderer, textures, targets, lr, loss="l2", betas=(0.9, 0.999), eps=1e-8,
                 optim_light=False, history=0, plane_stride=None, own_textures=False):
        self.renderer = renderer
        self.res, self.n_views = renderer.res, renderer.n_of_imgs
        dev = renderer.device
        if textures.device != dev or textures.dtype != torch.float32 or not textures.is_contiguous():
            raise ValueError(f"textures must be a contiguous float32 tensor on the renderer's device {dev}")
        if textures.numel() != 9 * self.res * self.res:
            raise ValueError(f"textures must hold 9x{self.res}x{self.res} values, got {tuple(textures.shape)}")
        if tuple(targets.shape) != (self.n_views, 3, self.res, self.res):
            raise ValueError(f"targets must be [{self.n_views}, 3, {self.res}, {self.res}], "
                             f"got {tuple(targets.shape)}")
        self.textures = textures
        self.lr = float(lr)
        self.beta1, self.beta2 = float(betas[0]), float(betas[1])
        self.eps = float(eps)
        self.mode = loss_mode(loss)
        self.optim_light = bool(optim_light)
        if self.optim_light:
            lp = renderer.light_pow
            if _lib.cuda_device(lp.device) != dev or lp.dtype != torch.float32 or not lp.is_contiguous() or lp.numel() != 3:
                raise ValueError("optim_light needs renderer.light_pow as a contiguous float32 [3] CUDA tensor")
            self.light_pow = lp
        else:
            self.light_pow = renderer.light_pow.detach().to(torch.float32).reshape(3).contiguous()
        # Adam's moments and the prepared targets belong to this object: they
        # are laid out with the library's padded plane stride (planes 4 MiB
        # apart at 1024^2 would share HBM channels) unless plane_stride says
        # otherwise (0: dense).  The textures are the caller's dense tensor,
        # updated in place, or with own_textures=True a copy at the same
        # padded stride that write_back() returns to the caller's tensor.
        if plane_stride is None:
            plane_stride = _lib.query("svbrdf_plane_stride", self.res)
        self.plane_stride = int(plane_stride)
        ps = self.plane_stride or self.res * self.res
        self.source = textures
        if own_textures:
            self._tex = torch.zeros(9 * ps, dtype=torch.float32, device=dev)
            self.textures = _planes(self._tex, textures.shape, ps)
            with torch.no_grad():
                self.textures.copy_(textures)
            self.texture_stride = self.plane_stride
        else:
            self.texture_stride = 0
        self._moments = torch.zeros((2, 9 * ps), dtype=torch.float32, device=dev)
        self.exp_avg = _planes(self._moments[0], textures.shape, ps)
        self.exp_avg_sq = _planes(self._moments[1], textures.shape, ps)
        self._targets = torch.zeros(3 * self.n_views * ps, dtype=torch.float32, device=dev)
        self.targets = _planes(self._targets, (self.n_views, 3, self.res, self.res), ps)
        tg = targets.detach().to(device=dev, dtype=torch.float32).contiguous()
        s = _lib.stream_handle(dev)
        dense = torch.empty_like(tg)
        _lib.call("svbrdf_prepare_targets", _lib.ptr(tg), tg.numel(), self.mode, _lib.ptr(dense), s)
        self.targets.copy_(dense)
        nstate = _lib.query("svbrdf_step_state_bytes")
        self.state = torch.zeros((nstate + 15) // 16 * 4, dtype=torch.float32, device=dev)
        nws = _lib.query("svbrdf_fused_step_workspace_bytes", self.res)
        # zero-filled once: it carries the arrival counter and the last
        # step's partial sums from one call to the next
        self.workspace = torch.zeros((nws + 15) // 16 * 4, dtype=torch.float32, device=dev)
        self.history = torch.zeros(max(int(history), 0), dtype=torch.float32, device=dev)
        _lib.call("svbrdf_step_state_init", _lib.ptr(self.state), self.beta1, self.beta2, s)
        self.steps_launched = 0
        self.graph = None
        self.graph_steps = 0
        self._graphs = {}

    def _launch(self, stream):
        r = self.renderer
        hist = self.history if self.history.numel() else None
        _lib.call("svbrdf_fused_step", _lib.ptr(self.textures), _lib.ptr(self.exp_avg), _lib.ptr(self.exp_avg_sq),
                  _lib.ptr(self.targets), self.res, self.texture_stride, self.plane_stride, self.n_views, r.size,
                  _lib.ptr(r.camera_pos), _lib.ptr(r.light_pos), _lib.ptr(self.light_pow), self.mode,
                  int(self.optim_light), self.lr, self.beta1, self.beta2, self.eps, _lib.ptr(self.state),
                  _lib.ptr(hist), self.history.numel(), _lib.ptr(self.workspace), self.workspace.numel() * 4,
                  stream)

    def step(self):
        """Launch one fused iteration on the current stream (asynchronous)."""
        self._launch(_lib.stream_handle(self.textures.device))
        self.steps_launched += 1

    def capture(self, steps=1):
        """Record `steps` consecutive iterations into one HIP graph
        (torch.cuda.CUDAGraph); every replay() then runs that many."""
        self.graph = self._capture_graph(steps)
        self.graph_steps = steps
        return self.graph

    def run(self, n, graph_steps=100):
        """n iterations as replays of cached graphs of `graph_steps`
        iterations (and one of the remainder); graph_steps=0 launches them
        one by one.  Each graph launch costs the GPU some tens of
        microseconds, so long loops go in large pieces."""
        if n <= 0:
            return
        if not graph_steps:
            for _ in range(n):
                self.step()
            return
        per = max(1, min(int(graph_steps), int(n)))
        for size in {per, n % per} - {0}:
            if size not in self._graphs:
                self._graphs[size] = self._capture_graph(size)
        for _ in range(n // per):
            self._graphs[per].replay()
        if n % per:
            self._graphs[n % per].replay()
        self.steps_launched += n

    def _capture_graph(self, steps):
        if steps < 1:
            raise ValueError("steps must be >= 1")
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=self.textures.device)
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            g.capture_begin()
            for _ in range(steps):
                self._launch(ctypes.c_void_p(s.cuda_stream))
            g.capture_end()
        torch.cuda.current_stream().wait_stream(s)
        return g

    def replay(self):
        self.graph.replay()
        self.steps_launched += self.graph_steps

    def write_back(self):
        """Copies the current maps into the tensor given at construction when
        this step keeps its own copy (own_textures=True); asynchronous."""
        if self.textures is not self.source:
            with torch.no_grad():
                self.source.copy_(self.textures)

    def flush(self):
        """Records the last step's loss if it is still pending (asynchronous).

        A step launch reduces the previous step's partial sums, so the loss of
        the most recent step needs this one-block launch before it is read."""
        hist = self.history if self.history.numel() else None
        _lib.call("svbrdf_fused_step_flush", _lib.ptr(self.state), self.res, self.n_views, _lib.ptr(hist),
                  self.history.numel(), _lib.ptr(self.workspace), self.workspace.numel() * 4,
                  _lib.stream_handle(self.textures.device))

    def read_state(self):
        """Host copy of the device step state, losses flushed (synchronises)."""
        self.flush()
        raw = self.state.detach().cpu().numpy().view(np.uint8)[:_STATE_DTYPE.itemsize]
        return raw.view(_STATE_DTYPE)[0]

    def loss(self):
        """Loss evaluated by the last completed step (synchronises)."""
        return float(self.read_state()["loss"])

    def losses(self):
        """Per-step losses recorded so far (synchronises)."""
        self.flush()
        n = min(self.steps_launched, self.history.numel())
        return self.history[:n].cpu().numpy()

    def state_dict(self):
        """Exact resume point: maps, Adam moments and the step state (flushed,
        so it does not refer to partial sums held in this workspace)."""
        self.flush()
        return {"textures": self.textures.detach().clone(), "exp_avg": self.exp_avg.clone(),
                "exp_avg_sq": self.exp_avg_sq.clone(), "state": self.state.clone(),
                "light_pow": self.light_pow.detach().clone()}

    def load_state_dict(self, sd):
        with torch.no_grad():
            self.textures.copy_(sd["textures"])
            self.exp_avg.copy_(sd["exp_avg"])
            self.exp_avg_sq.copy_(sd["exp_avg_sq"])
            self.state.copy_(sd["state"])
            # partial sums do not travel with a state: nothing is pending here
            self.state.view(torch.int32)[_STATE_DTYPE.fields["pending_step"][1] // 4] = 0
            if self.optim_light:
                self.light_pow.copy_(sd["light_pow"])
        raw = sd["state"].detach().cpu().numpy().view(np.uint8)[:_STATE_DTYPE.itemsize]
        self.steps_launched = int(raw.view(_STATE_DTYPE)[0]["step"])


class SvbrdfOptim:
    """src/svbrdf.py:15-83 (and the Optim base, src/optimization.py:10-42)."""

    def __init__(self, device, renderer_obj):
        self.device = torch.device(device)
        self.eps = 1e-4
        self.renderer_obj = renderer_obj
        self.res = renderer_obj.res
        self.loss = "l2"
        self.fused = None

    # --- initialisers: src/svbrdf.py:20-39 (same encodings, t = 2u - 1) ---
    def init_from_tex(self, textures):
        self.textures = textures.detach().to(self.device, torch.float32).contiguous().requires_grad_(True)

    def init_from_const(self, dif=0.5, spe=0.04, rgh=0.2):
        R, d = self.res, self.device
        normal = torch.zeros(1, 2, R, R, device=d)
        diffuse = torch.ones(1, 3, R, R, device=d) * dif * 2 - 1
        specular = torch.ones(1, 3, R, R, device=d) * spe * 2 - 1
        roughness = torch.ones(1, 1, R, R, device=d) * rgh * 2 - 1
        self.init_from_tex(torch.cat((diffuse, normal, roughness, specular), 1))

    def init_from_randn(self, dif=0.5, spe=0.04, rgh=0.2):
        R, d = self.res, self.device
        normal = (torch.randn(1, 2, R, R, device=d) / 4).clamp(-1, 1)
        diffuse = (torch.randn(1, 3, R, R, device=d) / 8 + dif).clamp(0, 1) * 2 - 1
        specular = (torch.randn(1, 3, R, R, device=d) / 32 + spe).clamp(0, 1) * 2 - 1
        roughness = (torch.randn(1, 1, R, R, device=d) / 16 + rgh).clamp(0, 1) * 2 - 1
        self.init_from_tex(torch.cat((diffuse, normal, roughness, specular), 1))

    def load_targets(self, targets):
        self.targets = targets

    def compute_image_loss(self, predicts):
        """src/optimization.py:28-29 (MSE); "log_l2" when self.loss says so."""
        if loss_mode(self.loss) == _lib.LOSS_L2:
            return torch.nn.functional.mse_loss(predicts, self.targets)
        floor = 1e-6 ** (1 / 2.2)
        return torch.nn.functional.mse_loss(torch.log(predicts), torch.log(self.targets.clamp(floor, 1)))

    def optim(self, epochs, lr, svbrdf_obj=None, optim_light=False, loss=None, save_every=100,
              sync_every=0, graph_steps=100):
        """src/svbrdf.py:44-83.

        The loop body runs as one fused kernel per epoch; the epochs between
        two host events (a save, a sync) run as replays of HIP graphs of
        ``graph_steps`` epochs.  Side outputs are as in the reference when
        ``svbrdf_obj`` provides them (maps and re-renders saved at epoch 1,
        every ``save_every`` epochs and the last epoch under
        optimize_dir/tmp/<timestamp>/<epoch>/).  Losses are recorded on the
        device and read back at the end (the host also synchronises every
        ``sync_every`` epochs when that is set), instead of with one
        ``.item()`` per epoch.
        """
        if loss is not None:
            self.loss = loss
        r = self.renderer_obj
        if optim_light:
            if svbrdf_obj is not None:
                lp = svbrdf_obj.cl[2]
            else:
                lp = r.light_pow
            lp = lp.detach().to(self.device, torch.float32).reshape(3).contiguous()
            if svbrdf_obj is not None:
                svbrdf_obj.cl[2] = lp
            r.update_light(lp)
        with torch.no_grad():
            tex = self.textures.detach()
        # the step keeps its own padded copy of the maps; self.textures is
        # brought up to date after every stretch of epochs
        self.fused = FusedStep(r, tex, self.targets, lr, loss=self.loss, optim_light=optim_light,
                               history=epochs, own_textures=True)
        tmp_dir = None
        if svbrdf_obj is not None and hasattr(svbrdf_obj, "optimize_dir"):
            stamp = str(datetime.datetime.now()).replace(" ", "-").replace(":", "-").replace(".", "-")
            tmp_dir = pathlib.Path(svbrdf_obj.optimize_dir) / "tmp" / stamp
            tmp_dir.mkdir(parents=True, exist_ok=True)
        for first, last, save in epoch_segments(epochs, save_every, sync_every, tmp_dir is not None):
            self.fused.run(last - first + 1, graph_steps)
            self.fused.write_back()
            if sync_every:
                torch.cuda.synchronize(self.device)
            if save:
                this_dir = tmp_dir / f"{last + 1}"
                this_dir.mkdir(parents=True, exist_ok=True)
                self.save_loss([self.fused.losses()], ["image loss"], tmp_dir / "loss.jpg", epochs)
                with torch.no_grad():
                    tc = self.textures.detach().clamp(-1, 1)
                    if hasattr(svbrdf_obj, "save_textures_th"):
                        svbrdf_obj.save_textures_th(tc, this_dir)
                    if hasattr(svbrdf_obj, "save_images_th"):
                        svbrdf_obj.save_images_th(r.eval(tc), this_dir)
        self.loss_image_list = [float(x) for x in self.fused.losses()]
        return self.loss_image_list

    def gradient(self, parameters):
        """src/optimization.py:17-23: the tensor (or each tensor of a list)
        with requires_grad set."""
        if isinstance(parameters, list):
            return [p.detach().requires_grad_(True) for p in parameters]
        return parameters.detach().requires_grad_(True)

    @staticmethod
    def save_loss(losses, labels, save_dir, N):
        """src/optimization.py:35-42: log(1 + loss) curves over N epochs.  Drawn
        on a standalone Agg figure, so pyplot's global state and the caller's
        backend are left alone."""
        from matplotlib.backends.backend_agg import FigureCanvasAgg
        from matplotlib.figure import Figure
        fig = Figure(figsize=(8, 4))
        FigureCanvasAgg(fig)
        ax = fig.add_subplot()
        for loss, label in zip(losses, labels):
            ax.plot(np.log1p(np.asarray(loss, dtype=np.float64)), label=label)
        ax.set_xlim(0, N)
        ax.legend()
        ax.set_title("log(1+loss)")
        fig.savefig(save_dir)
