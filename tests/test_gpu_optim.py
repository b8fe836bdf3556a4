"""The drop-in optimisation loop SvbrdfOptim.optim (src/svbrdf.py:44-83) and
the driver functions of src/scripts.py on top of it.

optim runs the epochs between two host events (a save, a sync) as replays of
HIP graphs of the fused step.  The replays launch the same kernel on the same
buffers as one launch per epoch, so the textures, Adam moments, light power
and per-epoch losses must be bit-identical to an eager FusedStep sequence.
Saves happen after epoch 1, every 100th and the last (src/svbrdf.py:74), hold
the textures of that epoch, and the loss recorded for the next epoch is the
image loss of exactly those textures (checked against torch's MSE of a
render, FP32 reduction order aside: rtol 1e-5)."""
import json
import pathlib

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


class Recorder:
    """The svbrdf_obj surface optim uses: optimize_dir, cl and the two savers."""

    def __init__(self, optimize_dir, cl):
        self.optimize_dir = optimize_dir
        self.cl = cl
        self.saved = []

    def save_textures_th(self, textures, d):
        self.saved.append(("tex", pathlib.Path(d).name, textures.detach().clone()))

    def save_images_th(self, images, d):
        self.saved.append(("img", pathlib.Path(d).name, images.detach().clone()))


def problem(pkg, res=32, V=9):
    syn = pkg.synthetic
    gt = syn.ground_truth_textures(res, seed=0)
    init = syn.init_textures(gt, seed=1)
    cam, light = syn.flash_grid(n_views=V)
    P = syn.light_power() * np.float32(0.8)        # so that optim_light has something to find
    cl = [torch.tensor(cam, device=DEV), torch.tensor(light, device=DEV), torch.tensor(P, device=DEV)]
    r = pkg.Microfacet(res, V, syn.PLANE_SIZE_CM, [c.clone() for c in cl], DEV)
    r.update_light(torch.tensor(syn.light_power(), device=DEV))
    with torch.no_grad():
        targets = r.eval(torch.tensor(gt, device=DEV)[None])
    return init, cl, targets


@pytest.mark.parametrize("graph_steps", [100, 7, 0])
@pytest.mark.parametrize("light", [False, True])
def test_optim_loop_is_the_fused_step_sequence(pkg, tmp_path, light, graph_steps):
    res, V, epochs, lr = 32, 9, 250, 0.01
    init, cl, targets = problem(pkg, res, V)
    size = pkg.synthetic.PLANE_SIZE_CM
    print(f"optim loop light={light} graph_steps={graph_steps}", flush=True)

    r1 = pkg.Microfacet(res, V, size, [c.clone() for c in cl], DEV)
    opt = pkg.SvbrdfOptim(DEV, r1)
    opt.init_from_tex(torch.tensor(init, device=DEV)[None])
    opt.load_targets(targets)
    rec = Recorder(tmp_path, [c.clone() for c in cl])
    losses = np.array(opt.optim(epochs, lr, rec, optim_light=light, graph_steps=graph_steps), np.float32)

    r2 = pkg.Microfacet(res, V, size, [c.clone() for c in cl], DEV)
    r2.update_light(cl[2].clone().contiguous())
    tex2 = torch.tensor(init, device=DEV)[None].contiguous()
    fs = pkg.FusedStep(r2, tex2, targets, lr, optim_light=light, history=epochs)
    snaps = {}
    for e in range(1, epochs + 1):
        fs.step()
        if e in (1, 100, 200, 250):
            snaps[e] = (tex2.clone(), r2.light_pow.clone())
    torch.cuda.synchronize()

    assert losses.shape == (epochs,) and np.all(np.isfinite(losses))
    np.testing.assert_array_equal(losses, fs.losses())
    assert torch.equal(opt.textures.detach(), tex2)
    assert torch.equal(opt.fused.exp_avg, fs.exp_avg) and torch.equal(opt.fused.exp_avg_sq, fs.exp_avg_sq)
    if light:
        assert torch.equal(rec.cl[2], r2.light_pow) and torch.equal(r1.light_pow, r2.light_pow)
        assert not torch.equal(rec.cl[2], cl[2])
    else:
        assert torch.equal(r1.light_pow, cl[2])

    # saves: after epochs 1, 100, 200 and 250, textures and re-renders of that epoch
    assert [n for k, n, _ in rec.saved if k == "tex"] == ["1", "100", "200", "250"]
    assert [n for k, n, _ in rec.saved if k == "img"] == ["1", "100", "200", "250"]
    imgs = {n: t for k, n, t in rec.saved if k == "img"}
    for k, n, t in rec.saved:
        if k != "tex":
            continue
        snap, lp = snaps[int(n)]
        assert torch.equal(t, snap.clamp(-1, 1)), n
        r2.update_light(lp)
        with torch.no_grad():
            img = r2.eval(snap.clamp(-1, 1))
        assert torch.equal(imgs[n], img), n
        if int(n) < epochs:
            # the next epoch's recorded loss is the image loss of these textures
            want = torch.nn.functional.mse_loss(img, targets).item()
            assert abs(losses[int(n)] - want) <= 1e-5 * want, (n, losses[int(n)], want)
    stamps = list((tmp_path / "tmp").iterdir())
    assert len(stamps) == 1 and (stamps[0] / "loss.jpg").is_file()
    assert sorted(p.name for p in stamps[0].iterdir() if p.is_dir()) == ["1", "100", "200", "250"]


def test_optim_sync_points_and_resume_from_losses(pkg):
    """sync_every splits the run further; the trajectory does not change."""
    res, V = 32, 9
    init, cl, targets = problem(pkg, res, V)
    out = []
    for sync_every, graph_steps in ((0, 100), (16, 5)):
        r = pkg.Microfacet(res, V, pkg.synthetic.PLANE_SIZE_CM, [c.clone() for c in cl], DEV)
        opt = pkg.SvbrdfOptim(DEV, r)
        opt.init_from_tex(torch.tensor(init, device=DEV)[None])
        opt.load_targets(targets)
        losses = opt.optim(90, 0.02, None, loss="log_l2", sync_every=sync_every, graph_steps=graph_steps)
        out.append((np.array(losses), opt.textures.detach().clone()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1])
    assert out[0][0][-1] < out[0][0][0]


def _dataset(pkg, tmp_path, res):
    """A JSON configuration with 9 flash views and reference maps saved as PNG."""
    syn = pkg.synthetic
    cam, light = syn.flash_grid(n_views=9)
    cfg = tmp_path / "cfg.json"
    cfg.write_text(json.dumps({
        "reference_dir": "ref", "target_dir": "target", "optimize_dir": "opt", "rerender_dir": "rr",
        "im_size": float(syn.PLANE_SIZE_CM), "idx": list(range(9)), "camera_pos": cam.tolist(),
        "light_pos": light.tolist(), "light_pow": syn.light_power().tolist()}))
    sio = pkg.SvbrdfIO(cfg, DEV)
    sio.save_textures_th(torch.tensor(syn.ground_truth_textures(res, seed=0), device=DEV)[None], sio.reference_dir)
    return cfg, sio


def test_scripts_render_then_optim_perpixel(pkg, tmp_path):
    from svbrdf_amd import imageio as io
    res = 64
    cfg, sio = _dataset(pkg, tmp_path, res)
    rendered = pkg.scripts.render(cfg, res)
    assert rendered.shape == (9, 3, res, res)
    for i in range(9):
        assert io.read_png(sio.target_dir / f"{i:02d}.png").shape == (res, res, 3)
    # starting from the maps the targets were rendered from, only 8-bit
    # quantisation separates them
    opt = pkg.scripts.optim_perpixel(cfg, res, 0.01, 3, "textures")
    assert opt.loss_image_list[0] < 1e-4
    opt = pkg.scripts.optim_perpixel(cfg, res, 0.01, 120, "const")
    li = opt.loss_image_list
    assert len(li) == 120 and li[-1] < 0.5 * li[0]
    for name in ("dif.png", "nom.png", "spe.png", "rgh.png", "tex.png"):
        assert (sio.optimize_dir / name).is_file()
    for i in range(9):
        assert (sio.rerender_dir / f"{i:02d}.png").is_file()
    runs = sorted((sio.optimize_dir / "tmp").iterdir())
    assert len(runs) == 2
    assert sorted(p.name for p in runs[-1].iterdir() if p.is_dir()) == ["1", "100", "120"]
    assert (runs[-1] / "120" / "tex.png").is_file() and (runs[-1] / "loss.jpg").is_file()
    # coarse-to-fine (run.py:55-56): the next stage reads this stage's saved
    # maps at twice the resolution (Lanczos-4 resize of maps and targets)
    d = json.loads(cfg.read_text())
    d.update(reference_dir="opt", optimize_dir="opt2", rerender_dir="rr2")
    (tmp_path / "cfg2.json").write_text(json.dumps(d))
    opt2 = pkg.scripts.optim_perpixel(tmp_path / "cfg2.json", 2 * res, 0.01, 5, "textures")
    assert opt2.textures.shape == (1, 9, 2 * res, 2 * res)
    assert opt2.loss_image_list[0] < 0.5 * li[0]
    assert (tmp_path / "opt2" / "tex.png").is_file()
    with pytest.raises(ValueError):
        pkg.scripts.optim_perpixel(cfg, res, 0.01, 1, "bogus")


def test_scripts_materialgan_drivers(pkg, tmp_path):
    res = 256
    cfg, sio = _dataset(pkg, tmp_path, res)
    pkg.scripts.render(cfg, res)
    tex = pkg.scripts.gen_textures_from_materialgan(cfg, res, seed=3)
    assert tex.shape == (1, 9, res, res) and (sio.reference_dir / "tex.png").is_file()
    opt = pkg.scripts.optim_ganlatent(cfg, res, 0.02, [3, 2, 1], [])
    assert len(opt.loss_image_list) == 3 and np.isfinite(opt.loss)
    with torch.no_grad():
        assert torch.allclose(opt.textures, opt.latent_to_textures(opt.latent), rtol=0, atol=1e-5)
    run = next((sio.optimize_dir / "tmp").iterdir())
    assert sorted(p.name for p in run.iterdir() if p.is_dir()) == ["1", "3"]
    latent = torch.load(run / "3" / "optim_latent.pt", weights_only=True)
    assert torch.equal(latent.to(DEV), opt.latent.detach())
    assert len(torch.load(run / "3" / "optim_noise.pt", weights_only=True)) == len(opt.noises)
    assert (sio.optimize_dir / "tex.png").is_file() and (sio.rerender_dir / "00.png").is_file()


def test_device_without_index(pkg, tmp_path):
    """Microfacet(..., "cuda") -- the reference's callers pass a device such
    as "cuda" or th.device("cuda:0") -- holds the resolved device, so the
    optimiser, the light-power path and the batch accept its tensors."""
    res, V = 16, 9
    init, cl, targets = problem(pkg, res, V)
    r = pkg.Microfacet(res, V, pkg.synthetic.PLANE_SIZE_CM, [c.clone() for c in cl], "cuda")
    assert r.device == torch.device("cuda", torch.cuda.current_device())
    opt = pkg.SvbrdfOptim("cuda", r)
    opt.init_from_tex(torch.tensor(init, device="cuda")[None])
    opt.load_targets(targets)
    losses = opt.optim(5, 0.01, Recorder(tmp_path, [c.clone() for c in cl]), optim_light=True)
    assert len(losses) == 5 and np.all(np.isfinite(losses))
    b = pkg.BatchedStep([(r, torch.tensor(init, device="cuda")[None].contiguous(), targets)], 0.01)
    b.step()
    torch.cuda.synchronize()
    assert np.isfinite(b.losses()).all()
