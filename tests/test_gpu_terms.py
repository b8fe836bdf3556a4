"""The reference's alternate BRDF terms on the GPU: Beckmann
(src/microfacet.py:34-38) and Fresnel_f0 (:40-41, f0 = the specular map) in
the forward and adjoint kernels (svbrdf_render_terms /
svbrdf_render_backward_terms, selected by Microfacet(..., ndf=, fresnel=)),
against the NumPy oracle's FP32 and FP64 renders and adjoints with the same
terms (the oracle's variants are pinned in tests/test_oracle.py by the NDF
normalisation integrals, closed forms and central differences).  Criteria:
tests/parity.py, as for the eval terms."""
import numpy as np
import pytest
import torch

from oracle import microfacet_np as M
from tests import parity
from tests.test_gpu_render import scene

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
VARIANTS = [("beckmann", "sg"), ("ggx", "schlick"), ("beckmann", "schlick")]


def gpu_render_terms(pkg, tex, res, size, cam, light, P, gout, ndf, fresnel):
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    r = pkg.Microfacet(res, cam.shape[0], size, cl, DEV, ndf=ndf, fresnel=fresnel)
    t = torch.tensor(tex, device=DEV)[None].requires_grad_(True)
    p = torch.tensor(P, device=DEV).requires_grad_(True)
    r.update_light(p)
    out = r.eval(t)
    (out * torch.tensor(gout, device=DEV)).sum().backward()
    torch.cuda.synchronize()
    return out.detach().cpu().numpy(), t.grad[0].cpu().numpy(), p.grad.cpu().numpy()


@pytest.mark.parametrize("ndf,fresnel", VARIANTS)
@pytest.mark.parametrize("res,V,coloc", [(16, 9, True), (33, 5, False), (64, 9, True)])
def test_variant_forward_and_adjoint(pkg, ndf, fresnel, res, V, coloc):
    tex, res, size, cam, light, P = scene(pkg, res, V, seed=res + V, coloc=coloc)
    gout = np.random.default_rng(res).standard_normal((V, 3, res, res)).astype(np.float32)
    out, dtex, dP = gpu_render_terms(pkg, tex, res, size, cam, light, P, gout, ndf, fresnel)
    kw = dict(ndf=ndf, fresnel=fresnel)
    o32 = M.render(tex, res, size, cam, light, P, np.float32, **kw)
    o64 = M.render(tex, res, size, cam, light, P, np.float64, **kw)
    sigma = M.fp32_noise_scale(tex, res, size, cam, light, P, **kw)
    name = f"{ndf}/{fresnel} R{res} V{V}"
    parity.check_forward(out, o32, o64, sigma, name)
    g32, _ = M.render_vjp(tex, res, size, cam, light, P, gout, np.float32, **kw)
    g64, p64 = M.render_vjp(tex, res, size, cam, light, P, gout, np.float64, **kw)
    parity.check_gradient(dtex, g32, g64, name)
    np.testing.assert_allclose(dP, p64, rtol=1e-4)


def test_variants_differ_and_eval_terms_are_the_default(pkg):
    tex, res, size, cam, light, P = scene(pkg, 16, 9, seed=3)
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    t = torch.tensor(tex, device=DEV)[None]
    outs = {}
    for ndf, fresnel in [("ggx", "sg")] + VARIANTS:
        r = pkg.Microfacet(res, 9, size, cl, DEV, ndf=ndf, fresnel=fresnel)
        with torch.no_grad():
            outs[(ndf, fresnel)] = r.eval(t).cpu().numpy()
    with torch.no_grad():
        default = pkg.Microfacet(res, 9, size, cl, DEV).eval(t).cpu().numpy()
    np.testing.assert_array_equal(default, outs[("ggx", "sg")])
    for v in VARIANTS:
        # near-normal incidence the two Fresnel forms differ by ~1e-4 of the
        # specular term only; Beckmann changes the highlight visibly
        assert np.abs(outs[v] - default).max() > (1e-3 if v[0] == "beckmann" else 0.0), v
