"""Drop-in microfacet renderer: the class surface of src/microfacet.py:9-120
backed by the gfx950 forward and adjoint kernels.

``Microfacet(res, n, size, cl, device)`` keeps the reference constructor,
the attributes its callers read (``res``, ``n_of_imgs``, ``eps``, ``f0``)
and the two methods they call (``eval``, ``update_light``).  ``eval`` is
differentiable with respect to the textures and the light power through
:class:`RenderFunction`, whose backward is the hand-written adjoint kernel.
"""
import torch

from . import _lib


def _as_device_f32(t, device, name):
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(t, dtype=torch.float32)
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {t.dtype}")
    return t.to(device).contiguous()


def _check_cuda(t, name):
    if t.device.type != "cuda":
        raise RuntimeError(f"{name} is on {t.device}: the HIP renderer runs on a GPU only "
                           "(there is no CPU path)")


class RenderFunction(torch.autograd.Function):
    """images = render(textures[1,9,R,R], light_pow[3]); backward = adjoint kernel."""

    @staticmethod
    def forward(ctx, textures, light_pow, renderer):
        _check_cuda(textures, "textures")
        tex = textures.detach().contiguous()
        if tex.dtype != torch.float32:
            raise TypeError(f"textures must be float32, got {tex.dtype}")
        lp = light_pow.detach().to(torch.float32).contiguous()
        res, V = renderer.res, renderer.n_of_imgs
        out = torch.empty((V, 3, res, res), dtype=torch.float32, device=tex.device)
        _lib.call("svbrdf_render", _lib.ptr(tex), res, V, renderer.size, _lib.ptr(renderer.camera_pos),
                  _lib.ptr(renderer.light_pos), _lib.ptr(lp), _lib.ptr(out), _lib.stream_handle(tex.device))
        ctx.save_for_backward(tex, lp)
        ctx.renderer = renderer
        ctx.tex_shape = textures.shape
        return out

    @staticmethod
    def backward(ctx, grad_out):
        tex, lp = ctx.saved_tensors
        r = ctx.renderer
        go = grad_out.detach().to(torch.float32).contiguous()
        gtex = torch.empty_like(tex)
        glp = None
        ws = None
        if ctx.needs_input_grad[1]:
            glp = torch.empty(3, dtype=torch.float32, device=tex.device)
            nbytes = _lib.query("svbrdf_render_backward_workspace_bytes", r.res)
            ws = torch.empty(nbytes // 4, dtype=torch.float32, device=tex.device)
        _lib.call("svbrdf_render_backward", _lib.ptr(tex), r.res, r.n_of_imgs, r.size, _lib.ptr(r.camera_pos),
                  _lib.ptr(r.light_pos), _lib.ptr(lp), _lib.ptr(go), _lib.ptr(gtex), _lib.ptr(glp),
                  _lib.ptr(ws), 0 if ws is None else ws.numel() * 4, _lib.stream_handle(tex.device))
        gtex = gtex.view(ctx.tex_shape) if ctx.needs_input_grad[0] else None
        if glp is not None:
            glp = glp.to(ctx.saved_tensors[1].dtype)
        return gtex, glp, None


class Microfacet:
    """src/microfacet.py:9-120 on the GPU.

    cl = [camera_pos [n,3], light_pos [n,3], light_pow [3]] in cm / per-RGB
    power, as built by SvbrdfIO.load_calibration_th (src/svbrdf.py:134-148).
    """

    def __init__(self, res, n, size, cl, device):
        self.res = int(res)
        self.n_of_imgs = int(n)
        self.f0 = 0.04
        self.eps = 1e-6
        self.size = float(size)
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError(f"Microfacet on {self.device}: the HIP renderer runs on a GPU only")
        _lib.load()
        self.camera_pos = _as_device_f32(cl[0], self.device, "camera_pos")
        self.light_pos = _as_device_f32(cl[1], self.device, "light_pos")
        if self.camera_pos.shape != (self.n_of_imgs, 3) or self.light_pos.shape != (self.n_of_imgs, 3):
            raise ValueError(f"camera_pos/light_pos must be [{self.n_of_imgs}, 3], got "
                             f"{tuple(self.camera_pos.shape)} / {tuple(self.light_pos.shape)}")
        self.update_light(cl[2])

    def update_light(self, light_pow):
        """src/microfacet.py:81-82.  Keeps the tensor (and its autograd graph)."""
        if not isinstance(light_pow, torch.Tensor):
            light_pow = torch.as_tensor(light_pow, dtype=torch.float32)
        if light_pow.device != self.device:
            light_pow = light_pow.to(self.device)
        if light_pow.numel() != 3:
            raise ValueError(f"light_pow must have 3 elements, got {light_pow.numel()}")
        self.light_pow = light_pow

    def eval(self, textures):
        """src/microfacet.py:84-120: textures [1,9,R,R] -> images [n,3,R,R]."""
        assert textures.shape[2] == textures.shape[3]
        assert self.res == textures.shape[2]
        if textures.dim() != 4 or textures.shape[0] != 1 or textures.shape[1] != 9:
            raise ValueError(f"textures must be [1, 9, R, R], got {tuple(textures.shape)}")
        return RenderFunction.apply(textures, self.light_pow.reshape(3), self)
