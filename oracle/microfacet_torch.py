"""Eager PyTorch restatement of the reference renderer at the reference's op
granularity.  TEST INFRASTRUCTURE ONLY (see oracle/__init__).

One torch op per term of src/microfacet.py:84-120, on full expanded
``[V,3,R,R]`` tensors, so that (a) torch.autograd gives the reference's own
gradient structure (clamp masks, expand-backward sums, the normalize
Jacobian) and (b) timing it on the CPU reproduces the cost structure of the
reference's CPU path (src/scripts.py:68 falls back to "cpu").

Pinned by: tests/golden/*.npz written by the reference's own Microfacet
(forward + vector-Jacobian products in FP32/FP64), gradcheck in FP64, and
the closed forms in tests/test_oracle.py.
"""
import math

import torch

EPS = 1e-6
LOSS_L2 = 0
LOSS_LOG_L2 = 1


class MicrofacetEager:
    """src/microfacet.py:9-120 restated.  Same constructor and methods."""

    def __init__(self, res, n, size, cl, device, dtype=torch.float32):
        self.res = res
        self.n_of_imgs = n
        self.eps = EPS
        self.dtype = dtype
        # src/microfacet.py:16-20 (the grid is built in float32 whatever the dtype)
        t = torch.arange(res, dtype=torch.float32, device=device)
        t = (((t + 0.5) / res - 0.5) * size).to(dtype)
        x, y = torch.meshgrid(t, t, indexing="xy")
        pos = torch.stack((x, -y, torch.zeros_like(x)), 2)
        self.pos = pos.permute(2, 0, 1).unsqueeze(0).expand(n, -1, -1, -1)
        # src/microfacet.py:22-24
        cam = cl[0].to(device=device, dtype=dtype)
        light = cl[1].to(device=device, dtype=dtype)
        self.camera_pos = cam[:, :, None, None].expand(-1, -1, res, res)
        self.light_pos = light[:, :, None, None].expand(-1, -1, res, res)
        self.update_light(cl[2].to(device=device, dtype=dtype))

    def update_light(self, light_pow):
        # src/microfacet.py:81-82
        self.light_pow = light_pow[None, :, None, None].expand(self.n_of_imgs, -1, self.res, self.res)

    @staticmethod
    def _dot(a, b):
        # src/microfacet.py:54-55
        return (a * b).sum(1, keepdim=True).expand(-1, 3, -1, -1)

    @staticmethod
    def _normalize(vec):
        # src/microfacet.py:57-58
        return vec / vec.norm(2.0, 1, keepdim=True)

    def _get_dir(self, p):
        # src/microfacet.py:60-62
        vec = p - self.pos
        return self._normalize(vec), self._dot(vec, vec)

    def _normal(self, t2):
        # src/microfacet.py:64-70
        nx = t2[:, 0].clamp(-1, 1)
        ny = t2[:, 1].clamp(-1, 1)
        s = (nx ** 2 + ny ** 2).clamp(0, 1 - self.eps)
        nz = (1 - s).sqrt()
        return self._normalize(torch.stack((nx, ny, nz), 1))

    def _maps(self, tex):
        # src/microfacet.py:72-79
        n = self.n_of_imgs
        dif = (((tex[:, 0:3] + 1) / 2) ** 2.2).expand(n, -1, -1, -1)
        nrm = self._normal(tex[:, 3:5]).expand(n, -1, -1, -1)
        rgh = (((tex[:, 5] + 1) / 2) ** 2.2).expand(n, 3, -1, -1)
        spe = (((tex[:, 6:9] + 1) / 2) ** 2.2).expand(n, -1, -1, -1)
        return nrm, dif, spe, rgh

    def _ggx(self, cos_h, alpha):
        # src/microfacet.py:28-32
        c2 = cos_h ** 2
        a2 = alpha ** 2
        den = c2 * a2 + (1 - c2)
        return a2 / (math.pi * den ** 2 + self.eps)

    @staticmethod
    def _fresnel(cos, spe):
        # src/microfacet.py:43-45
        sphg = torch.pow(2.0, ((-5.55473 * cos) - 6.98316) * cos)
        return spe + (1.0 - spe) * sphg

    def _smith(self, nv, nl, alpha):
        # src/microfacet.py:47-52
        k = alpha * 0.5 + self.eps
        g1v = nv / (nv * (1.0 - k) + k)
        g1l = nl / (nl * (1.0 - k) + k)
        return g1v * g1l

    def eval(self, textures):
        # src/microfacet.py:84-120
        assert textures.shape[2] == textures.shape[3]
        assert self.res == textures.shape[2]
        return self.radiance(textures).clamp(self.eps, 1) ** (1 / 2.2)

    def radiance(self, textures):
        # src/microfacet.py:86-117: the image before the output clamp and gamma
        nrm, dif, spe, rgh = self._maps(textures)
        v, _ = self._get_dir(self.camera_pos)
        l, d2 = self._get_dir(self.light_pos)
        h = self._normalize(l + v)
        nv = self._dot(nrm, v).clamp(min=0)
        nl = self._dot(nrm, l).clamp(min=0)
        nh = self._dot(nrm, h).clamp(min=0)
        vh = self._dot(v, h).clamp(min=0)
        f1 = dif / math.pi
        f1 = f1 * (1 - spe)
        D = self._ggx(nh, rgh ** 2)
        F = self._fresnel(vh, spe)
        G = self._smith(nv, nl, rgh ** 2)
        f2 = D * F * G / (4 * nv * nl + self.eps)
        f = f1 + f2
        return self.light_pow * f * nl / d2


def target_floor():
    return EPS ** (1 / 2.2)


def image_loss(out, target, mode):
    """L2 = torch.nn.MSELoss() (src/optimization.py:14,29); LOG_L2 = mean of
    (ln out - ln clamp(target, eps^(1/2.2), 1))^2 (this project's definition)."""
    if mode == LOSS_L2:
        return torch.nn.functional.mse_loss(out, target)
    tl = torch.log(target.clamp(target_floor(), 1))
    return torch.nn.functional.mse_loss(torch.log(out), tl)


def optimise(renderer, textures, targets, epochs, lr, mode, optim_light=False,
             light_pow=None, foreach=None, record=None):
    """The loop of src/svbrdf.py:44-71 without its PNG/plot side branch
    (:74-83): Adam(lr, betas=(0.9, 0.999)), clamp on read, loss, backward,
    step.  Returns (textures, light_pow, optimizer, losses)."""
    textures = textures.detach().clone().requires_grad_(True)
    params = [textures]
    if optim_light:
        light_pow = light_pow.detach().clone().requires_grad_(True)
        params.append(light_pow)
    opt = torch.optim.Adam(params, lr=lr, betas=(0.9, 0.999), foreach=foreach)
    losses = []
    for _ in range(epochs):
        if optim_light:
            renderer.update_light(light_pow)
        out = renderer.eval(textures.clamp(-1, 1))
        loss = image_loss(out, targets, mode)
        losses.append(loss.detach())
        opt.zero_grad()
        loss.backward()
        opt.step()
        if record is not None:
            record.append(textures.detach().clone())
    return textures.detach(), (light_pow.detach() if optim_light else None), opt, losses
