"""256^2 x 9 views: the render adjoint receives dL/dimages from torch's convolution
backward (a VGG-19-shaped feature loss).  The texture gradient through the
HIP renderer is compared with the same loss through the eager restatement of
the reference renderer (FP32 and FP64), under tests/parity.py's criteria."""
import numpy as np
import pytest
import torch

from oracle import microfacet_torch as MT
from tests import convnet_ref as R
from tests import parity

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def setup(pkg, res=32, V=9):
    gt = pkg.synthetic.ground_truth_textures(res, seed=0)
    init = pkg.synthetic.init_textures(gt, seed=1)
    cam, light = pkg.synthetic.flash_grid(n_views=V)
    P = pkg.synthetic.light_power()
    return gt, init, cam, light, P, pkg.synthetic.PLANE_SIZE_CM


class _TorchStack:
    """compose.VGGShapedFeatures' weights through F.conv2d in the input's dtype."""

    def __init__(self, net):
        self.net = net

    def __call__(self, x):
        return R.vgg_shaped_ref(self.net, x)


def feature_grad(pkg, renderer_cls, tex0, gt, cam, light, P, size, dtype):
    from svbrdf_amd import compose
    res = tex0.shape[-1]
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    net = compose.VGGShapedFeatures(seed=4, device=DEV)
    if renderer_cls is None:
        r = pkg.Microfacet(res, cam.shape[0], size, cl, DEV)
    else:
        # the reference side: eager renderer and the stack's torch restatement
        r = renderer_cls(res, cam.shape[0], size, cl, DEV, dtype)
        net = _TorchStack(net)
    with torch.no_grad():
        target = r.eval(torch.tensor(gt, device=DEV, dtype=dtype)[None])
    # the reference side's Gram matrices by torch autograd of bmm
    fl = compose.FeatureLoss(net, target, gram_fn=None if renderer_cls is None else R.gram_ref)
    t = torch.tensor(tex0, device=DEV, dtype=dtype)[None].requires_grad_(True)
    L = fl(r.eval(t.clamp(-1, 1)))
    L.backward()
    return L.item(), t.grad[0].double().cpu().numpy()


@pytest.mark.parametrize("res", [32, 256])
def test_feature_loss_gradient_through_hip_adjoint(pkg, res):
    """Config 4's workload shape (256^2 x 9 views, VGG-19-shaped stack to
    conv5_1, Gram L2 on 5 layers + pixel L2) and a small case."""
    gt, init, cam, light, P, size = setup(pkg, res=res)
    L, g = feature_grad(pkg, None, init, gt, cam, light, P, size, torch.float32)
    L32, g32 = feature_grad(pkg, MT.MicrofacetEager, init, gt, cam, light, P, size, torch.float32)
    L64, g64 = feature_grad(pkg, MT.MicrofacetEager, init, gt, cam, light, P, size, torch.float64)
    assert abs(L - L64) <= 1e-4 * abs(L64)
    errs = parity.check_gradient(g, g32, g64, f"feature loss {res}^2")
    print(f"feature loss {res}^2 x 9: per-channel |g - g64|/|g64| max {max(errs):.2e}")


def test_feature_loss_optimisation_runs(pkg):
    from svbrdf_amd import compose
    gt, init, cam, light, P, size = setup(pkg, res=64)
    cl = [torch.tensor(cam), torch.tensor(light), torch.tensor(P)]
    r = pkg.Microfacet(64, 9, size, cl, DEV)
    net = compose.VGGShapedFeatures(seed=4, device=DEV)
    with torch.no_grad():
        target = r.eval(torch.tensor(gt, device=DEV)[None])
    fl = compose.FeatureLoss(net, target)
    t = torch.tensor(init, device=DEV)[None].requires_grad_(True)
    opt = torch.optim.Adam([t], lr=0.01)
    losses = []
    for _ in range(10):
        L = fl(r.eval(t.clamp(-1, 1)))
        opt.zero_grad()
        L.backward()
        opt.step()
        losses.append(L.item())
    assert np.all(np.isfinite(losses)) and losses[-1] < losses[0]
