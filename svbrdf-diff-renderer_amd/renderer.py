"""Drop-in microfacet renderer: the class surface of src/microfacet.py:9-120
backed by the gfx950 forward and adjoint kernels.

``Microfacet(res, n, size, cl, device)`` keeps the reference constructor,
the attributes its callers read (``res``, ``n_of_imgs``, ``eps``, ``f0``)
and the two methods they call (``eval``, ``update_light``).  ``eval`` is
differentiable with respect to the textures and the light power through
:class:`RenderFunction`, whose backward is the hand-written adjoint kernel.

Keyword-only ``ndf="beckmann"`` / ``fresnel="schlick"`` render with the
reference's unused terms (Beckmann, microfacet.py:34-38; Fresnel_f0 with f0 =
the specular map, :40-41) in both kernels; the defaults are eval's.  The
public term methods of the reference class (GGX, Beckmann, Fresnel_f0,
Fresnel, Smith, dot, normalize, get_dir, reconstruct_normal, tex2map) are kept
as the same elementwise formulas on torch tensors for callers that use them
directly; eval does not go through them.
"""
import numpy as np
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
        _lib.call("svbrdf_render_terms", _lib.ptr(tex), res, V, renderer.size, _lib.ptr(renderer.camera_pos),
                  _lib.ptr(renderer.light_pos), _lib.ptr(lp), renderer.terms, _lib.ptr(out),
                  _lib.stream_handle(tex.device))
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
        _lib.call("svbrdf_render_backward_terms", _lib.ptr(tex), r.res, r.n_of_imgs, r.size,
                  _lib.ptr(r.camera_pos), _lib.ptr(r.light_pos), _lib.ptr(lp), r.terms, _lib.ptr(go),
                  _lib.ptr(gtex), _lib.ptr(glp), _lib.ptr(ws), 0 if ws is None else ws.numel() * 4,
                  _lib.stream_handle(tex.device))
        gtex = gtex.view(ctx.tex_shape) if ctx.needs_input_grad[0] else None
        if glp is not None:
            glp = glp.to(ctx.saved_tensors[1].dtype)
        return gtex, glp, None


class Microfacet:
    """src/microfacet.py:9-120 on the GPU.

    cl = [camera_pos [n,3], light_pos [n,3], light_pow [3]] in cm / per-RGB
    power, as built by SvbrdfIO.load_calibration_th (src/svbrdf.py:134-148).
    """

    NDFS = {"ggx": 0, "beckmann": _lib.NDF_BECKMANN}
    FRESNELS = {"sg": 0, "schlick": _lib.FRESNEL_SCHLICK}

    def __init__(self, res, n, size, cl, device, *, ndf="ggx", fresnel="sg"):
        if ndf not in self.NDFS or fresnel not in self.FRESNELS:
            raise ValueError(f"ndf must be one of {list(self.NDFS)}, fresnel one of {list(self.FRESNELS)}")
        self.ndf, self.fresnel = ndf, fresnel
        self.terms = self.NDFS[ndf] | self.FRESNELS[fresnel]
        self.res = int(res)
        self.n_of_imgs = int(n)
        self.f0 = 0.04
        self.eps = 1e-6
        self.size = float(size)
        self.device = _lib.cuda_device(device, "Microfacet")
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

    # ------------------------------------------------------------------
    # The reference's public term methods (src/microfacet.py:28-79), as the
    # same elementwise formulas on torch tensors.
    def GGX(self, cos_h, alpha):
        c2 = cos_h ** 2
        a2 = alpha ** 2
        den = c2 * a2 + (1 - c2)
        return a2 / (np.pi * den ** 2 + self.eps)

    def Beckmann(self, cos_h, alpha):
        c2 = cos_h ** 2
        t2 = (1 - c2) / c2
        a2 = alpha ** 2
        return torch.exp(-t2 / a2) / (np.pi * a2 * c2 ** 2)

    def Fresnel_f0(self, cos, f0):
        return f0 + (1 - f0) * (1 - cos) ** 5

    def Fresnel(self, cos, specular):
        sphg = torch.pow(2.0, ((-5.55473 * cos) - 6.98316) * cos)
        return specular + (1.0 - specular) * sphg

    def Smith(self, n_dot_v, n_dot_l, alpha):
        k = alpha * 0.5 + self.eps
        return (n_dot_v / (n_dot_v * (1.0 - k) + k)) * (n_dot_l / (n_dot_l * (1.0 - k) + k))

    def dot(self, a, b):
        return (a * b).sum(1, keepdim=True).expand(-1, 3, -1, -1)

    def normalize(self, vec):
        return vec / (vec.norm(2.0, 1, keepdim=True))

    @property
    def pos(self):
        """[1, 3, R, R] pixel centres (x, -y, 0) in cm (src/microfacet.py:16-19)."""
        c = ((torch.arange(self.res, dtype=torch.float32, device=self.device) + 0.5) / self.res - 0.5) * self.size
        yy, xx = torch.meshgrid(c, c, indexing="ij")
        return torch.stack((xx, -yy, torch.zeros_like(xx)))[None]

    def get_dir(self, pos):
        """src/microfacet.py:60-62: pos [n,3,R,R] (or [n,3]) -> unit direction
        [n,3,R,R] and squared distance (expanded to 3 channels)."""
        if pos.dim() == 2:
            pos = pos.reshape(-1, 3, 1, 1)
        vec = pos - self.pos
        return self.normalize(vec), self.dot(vec, vec)

    def reconstruct_normal(self, texture):
        """src/microfacet.py:64-70 on [N, 2, R, R]."""
        normal_x = texture[:, 0, :, :].clamp(-1, 1)
        normal_y = texture[:, 1, :, :].clamp(-1, 1)
        normal_xy = (normal_x ** 2 + normal_y ** 2).clamp(0, 1 - self.eps)
        normal_z = (1 - normal_xy).sqrt()
        return self.normalize(torch.stack((normal_x, normal_y, normal_z), 1))

    def tex2map(self, textures):
        """src/microfacet.py:72-79: textures [1,9,R,R] -> (normal, diffuse,
        specular, roughness), each expanded to [n,3,R,R]."""
        n = self.n_of_imgs
        diffuse = (((textures[:, 0:3] + 1) / 2) ** 2.2).expand(n, -1, -1, -1)
        normal = self.reconstruct_normal(textures[:, 3:5]).expand(n, -1, -1, -1)
        roughness = (((textures[:, 5] + 1) / 2) ** 2.2).expand(n, 3, -1, -1)
        specular = (((textures[:, 6:9] + 1) / 2) ** 2.2).expand(n, -1, -1, -1)
        return normal, diffuse, specular, roughness
