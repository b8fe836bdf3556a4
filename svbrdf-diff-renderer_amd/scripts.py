"""The reference's driver functions for the renderer path (src/scripts.py:
31-41, 67-97, 18-28, 100-138), on the HIP renderer and the fused optimiser.

Each takes the reference's arguments and reads / writes the same files
through :class:`SvbrdfIO` (JSON configuration, 8-bit PNG maps and views):

* :func:`render` -- maps from ``reference_dir`` rendered into ``target_dir``;
* :func:`optim_perpixel` -- per-pixel optimisation against ``target_dir``,
  maps to ``optimize_dir``, re-renders to ``rerender_dir``;
* :func:`gen_textures_from_materialgan` and :func:`optim_ganlatent` -- the
  MaterialGAN path, with the seeded stand-in networks of
  :mod:`svbrdf_amd.materialgan` (the pretrained checkpoints are downloads).

The device is always ``cuda:0``: there is no CPU path, so these raise when no
GPU is visible instead of falling back.  The capture and Mitsuba environment
map drivers (src/scripts.py:44-64) need libraries outside the renderer path
and are not provided.
"""
import torch

from .materialgan import MaterialGANOptim
from .optim import SvbrdfOptim
from .renderer import Microfacet
from .svbrdf_io import SvbrdfIO


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("the HIP renderer needs a GPU (cuda:0); none is visible")
    return torch.device("cuda:0")


def render(json_dir, res):
    """src/scripts.py:31-41."""
    device = _device()
    svbrdf_obj = SvbrdfIO(json_dir, device)
    textures = svbrdf_obj.load_textures_th(svbrdf_obj.reference_dir, res)
    render_obj = Microfacet(res, svbrdf_obj.n_of_imgs, svbrdf_obj.im_size, svbrdf_obj.cl, device)
    with torch.no_grad():
        rendereds = render_obj.eval(textures)
    svbrdf_obj.save_images_th(rendereds, svbrdf_obj.target_dir)
    return rendereds


def optim_perpixel(json_dir, res, lr, epochs, tex_init, optim_light=False, loss="l2"):
    """src/scripts.py:67-97.  tex_init: "random", "const" or "textures" (the
    maps in reference_dir).  ``loss`` selects the image loss of the fused
    step ("l2" as the reference, or "log_l2").  Returns the optimiser."""
    device = _device()
    svbrdf_obj = SvbrdfIO(json_dir, device)
    targets = svbrdf_obj.load_images_th(svbrdf_obj.target_dir, res)
    renderer_obj = Microfacet(res, svbrdf_obj.n_of_imgs, svbrdf_obj.im_size, svbrdf_obj.cl, device)
    optim_obj = SvbrdfOptim(device, renderer_obj)
    optim_obj.load_targets(targets)
    if tex_init == "random":
        optim_obj.init_from_randn()
    elif tex_init == "const":
        optim_obj.init_from_const()
    elif tex_init == "textures":
        optim_obj.init_from_tex(svbrdf_obj.load_textures_th(svbrdf_obj.reference_dir, res))
    else:
        raise ValueError(f"tex_init must be 'random', 'const' or 'textures', got {tex_init!r}")
    optim_obj.optim(epochs, lr, svbrdf_obj, optim_light, loss=loss)
    with torch.no_grad():
        final = optim_obj.textures.detach().clamp(-1, 1)
        svbrdf_obj.save_textures_th(final, svbrdf_obj.optimize_dir)
        if optim_light:
            print("Optimized light: ", svbrdf_obj.cl[2])
            renderer_obj.update_light(svbrdf_obj.cl[2])
        svbrdf_obj.save_images_th(renderer_obj.eval(final), svbrdf_obj.rerender_dir)
    return optim_obj


def gen_textures_from_materialgan(json_dir, res=256, seed=0):
    """src/scripts.py:18-28: maps synthesised from a random latent, saved to
    reference_dir."""
    device = _device()
    svbrdf_obj = SvbrdfIO(json_dir, device)
    optim_obj = MaterialGANOptim(device, _ResOnly(res))
    optim_obj.init_from([], seed=seed)
    with torch.no_grad():
        textures = optim_obj.latent_to_textures(optim_obj.latent)
    svbrdf_obj.save_textures_th(textures, svbrdf_obj.reference_dir)
    return textures


def optim_ganlatent(json_dir, res, lr, epochs, tex_init, optim_light=False):
    """src/scripts.py:100-138.  epochs = [total, latent epochs, noise epochs]
    per cycle; tex_init = [] (a random latent), [latent.pt] or [latent.pt,
    noise.pt], or "auto": two short runs from two starting latents, keeping
    the one with the lower image loss (the reference starts them from its
    average and constant W+ checkpoints; here they are seeded latents)."""
    device = _device()
    svbrdf_obj = SvbrdfIO(json_dir, device)
    targets = svbrdf_obj.load_images_th(svbrdf_obj.target_dir, res)
    renderer_obj = Microfacet(res, svbrdf_obj.n_of_imgs, svbrdf_obj.im_size, svbrdf_obj.cl, device)
    optim_obj = MaterialGANOptim(device, renderer_obj)
    optim_obj.load_targets(targets)
    if tex_init == "auto":
        optim_obj.init_from([], seed=0)
        optim_obj.optim([80, 10, 10], lr, svbrdf_obj, optim_light)
        loss_dif = optim_obj.loss_image
        optim_obj.init_from([], seed=1)
        optim_obj.optim([80, 10, 10], lr, svbrdf_obj, optim_light)
        loss_spe = optim_obj.loss_image
        if loss_dif < loss_spe:
            print("Non-specular material!")
            optim_obj.init_from([], seed=0)
        else:
            print("Specular material!")
            optim_obj.init_from([], seed=1)
    else:
        optim_obj.init_from(tex_init)
    optim_obj.optim(epochs, lr, svbrdf_obj, optim_light)
    with torch.no_grad():
        svbrdf_obj.save_textures_th(optim_obj.textures, svbrdf_obj.optimize_dir)
        if optim_light:
            print("Optimized light: ", svbrdf_obj.cl[2])
            renderer_obj.update_light(svbrdf_obj.cl[2])
        svbrdf_obj.save_images_th(renderer_obj.eval(optim_obj.textures), svbrdf_obj.rerender_dir)
    return optim_obj


class _ResOnly:
    """The generator is sized from its renderer; synthesis alone needs only the
    resolution."""

    def __init__(self, res):
        self.res = int(res)
