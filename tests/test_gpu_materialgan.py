"""MaterialGAN caller (src/materialgan.py:21-153) composed with the HIP render
adjoint.  The generator and VGG weights are seeded stand-ins for the
downloaded checkpoints, so their outputs have no reference to match; what is
checked is the composition: the gradient that reaches the textures through
the HIP adjoint under the reference's loss (image MSE + 0.1 x the VGG feature
loss of src/descriptor.py) against the same loss through the eager
restatement of the renderer in FP32 and FP64 (tests/parity.py criteria), and
that the alternating latent / noise loop runs and moves the right variables."""
import numpy as np
import pytest
import torch

from oracle import microfacet_torch as MT
from tests import convnet_ref as R
from tests import parity

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def scene(pkg, res=64, V=9):
    cam, light = pkg.synthetic.flash_grid(n_views=V)
    return cam, light, pkg.synthetic.light_power(), pkg.synthetic.PLANE_SIZE_CM


def test_generator_shapes(pkg):
    from svbrdf_amd import materialgan as MG
    g = MG.StyleGAN2Generator(256, 9, seed=1, device=DEV)
    z = torch.randn(1, 512, device=DEV)
    wp = g.truncation(g.mapping(z))
    assert wp.shape == (1, 14, 512)
    out = g.synthesis(wp, g.make_noises(torch.Generator().manual_seed(0)))
    assert out.shape == (1, 9, 256, 256) and torch.isfinite(out).all()


def _loss_grad(renderer, vgg, tex, targets, dtype):
    """Image MSE + 0.1 x the feature loss; the feature loss on the HIP
    convolutions (vgg itself, FP32) or on its torch restatement (any dtype)."""
    t = tex.detach().to(dtype).clone().requires_grad_(True)
    out = renderer.eval(t)
    if dtype == torch.float32 and renderer.__class__.__name__ == "Microfacet":
        feat = vgg(out)
    else:
        ref_t = R.vgg_features_ref(vgg, targets.to(dtype))
        feat = torch.nn.functional.mse_loss(R.vgg_features_ref(vgg, out), ref_t)
    loss = torch.nn.functional.mse_loss(out, targets.to(dtype)) + 0.1 * feat
    loss.backward()
    return float(loss), t.grad[0].double().cpu().numpy()


def test_feature_loss_gradient_through_hip_adjoint(pkg):
    from svbrdf_amd import materialgan as MG
    res = 64
    cam, light, P, size = scene(pkg, res)
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    gt = torch.tensor(pkg.synthetic.ground_truth_textures(res, seed=0), device=DEV)[None]
    tex = torch.tensor(pkg.synthetic.init_textures(gt[0].cpu().numpy(), seed=1), device=DEV)[None]
    hip = pkg.Microfacet(res, 9, size, cl, DEV)
    with torch.no_grad():
        targets = hip.eval(gt)
    vgg = MG.VGGLoss(DEV, np.array([1, 1, 4, 8]) / 14)
    vgg.load(targets)
    L, g = _loss_grad(hip, vgg, tex, targets, torch.float32)
    e32 = MT.MicrofacetEager(res, 9, size, cl, DEV, torch.float32)
    e64 = MT.MicrofacetEager(res, 9, size, cl, DEV, torch.float64)
    L32, g32 = _loss_grad(e32, vgg, tex, targets, torch.float32)
    L64, g64 = _loss_grad(e64, vgg, tex, targets, torch.float64)
    assert abs(L - L64) <= 1e-4 * abs(L64)
    parity.check_gradient(g, g32, g64, "materialgan loss")


def test_materialgan_loop_runs_and_alternates(pkg):
    res = 256
    cam, light, P, size = scene(pkg, res)
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    r = pkg.Microfacet(res, 9, size, cl, DEV)
    from svbrdf_amd import materialgan as MG
    # targets: renders of what a differently seeded latent synthesises
    opt_t = MG.MaterialGANOptim(DEV, r, seed=7)
    opt_t.init_from(seed=123)
    with torch.no_grad():
        targets = r.eval(opt_t.latent_to_textures(opt_t.latent))
    opt = MG.MaterialGANOptim(DEV, r, seed=7)
    opt.init_from(seed=5)
    opt.load_targets(targets)
    lat0 = opt.latent.detach().clone()
    noise0 = [n.detach().clone() for n in opt.noises]
    # 2 latent epochs: the noise must not move
    li, lf = opt.optim((2, 2, 1), 0.02)
    assert len(li) == 2 and np.all(np.isfinite(li)) and np.all(np.isfinite(lf))
    assert not torch.equal(opt.latent.detach(), lat0)
    assert all(torch.equal(n.detach(), m) for n, m in zip(opt.noises, noise0))
    # then a noise epoch: the latent must not move
    lat1 = opt.latent.detach().clone()
    opt.optim((3, 0, 1), 0.02)
    assert torch.equal(opt.latent.detach(), lat1)
    assert any(not torch.equal(n.detach(), m) for n, m in zip(opt.noises, noise0))
    assert opt.textures.shape == (1, 9, res, res)


def test_graph_epochs_match_eager_adam(pkg):
    """optim(graph=True) replays each epoch from a HIP graph per kind (the
    Adam step written out as the foreach ops torch.optim.Adam runs); from the
    same start it must give the eager loop's latent, noise maps and losses
    bit for bit, across latent -> noise -> latent switches and replays of both
    graphs, and twice in a row (a second optim() call reuses the graphs)."""
    res = 256
    cam, light, P, size = scene(pkg, res)
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    r = pkg.Microfacet(res, 9, size, cl, DEV)
    from svbrdf_amd import materialgan as MG
    opt_t = MG.MaterialGANOptim(DEV, r, seed=7)
    opt_t.init_from(seed=123)
    with torch.no_grad():
        targets = r.eval(opt_t.latent_to_textures(opt_t.latent))
    runs = []
    for graph in (False, True):
        opt = MG.MaterialGANOptim(DEV, r, seed=7)
        opt.init_from(seed=5)
        opt.load_targets(targets)
        li, lf = opt.optim((7, 2, 2), 0.02, graph=graph)
        li2, lf2 = opt.optim((3, 1, 1), 0.01, graph=graph)
        torch.cuda.synchronize()
        runs.append((opt.latent.detach().clone(), [n.detach().clone() for n in opt.noises], li + li2, lf + lf2,
                     opt.textures.clone()))
    (la, na, lia, lfa, ta), (lb, nb, lib, lfb, tb) = runs
    assert torch.equal(la, lb)
    assert all(torch.equal(x, y) for x, y in zip(na, nb))
    assert lia == lib and lfa == lfb
    assert torch.equal(ta, tb)
