"""MI355X-native microfacet renderer and per-pixel SVBRDF optimiser.

Drop-in for the hot path of tflsguoyu/svbrdf-diff-renderer:

* :class:`Microfacet` -- src/microfacet.py:9-120 (forward + adjoint kernels)
* :class:`SvbrdfOptim` -- src/svbrdf.py:15-83 (one fused kernel per iteration)
* :class:`FusedStep` -- the device-resident optimisation problem behind it
  (graph capture, exact resume)
* :class:`BatchedStep` -- many materials of uneven size in one persistent
  launch per iteration, with an autotuned launch shape
* :mod:`materialgan` -- the MaterialGAN caller (src/materialgan.py) on the
  HIP render adjoint, with seeded stand-ins for the downloaded networks
* :class:`SvbrdfIO`, :mod:`imageio` -- src/svbrdf.py:86-221 and
  src/imageio.py: maps and targets between 8-bit PNG files and tensors, the
  per-pixel codec and the Lanczos-4 resize on the GPU
* :mod:`scripts` -- the reference's driver functions (render,
  optim_perpixel, optim_ganlatent, gen_textures_from_materialgan,
  src/scripts.py) over JSON configurations and PNG files

The directory name is not a Python identifier; it is imported under the
module name ``svbrdf_amd`` (see ``load_package`` in __graft_entry__.py).
All compute runs in libsvbrdf_hip.so (gfx950); nothing falls back to the CPU.
"""
from . import _lib, imageio, materialgan, scripts, synthetic
from ._lib import LOSS_L2, LOSS_LOG_L2, SvbrdfHipError
from .batch import BatchedStep
from .optim import FusedStep, SvbrdfOptim, loss_mode, upsample_textures
from .renderer import Microfacet, RenderFunction
from .svbrdf_io import SvbrdfIO, reload_textures

__all__ = ["Microfacet", "RenderFunction", "SvbrdfOptim", "FusedStep", "BatchedStep", "SvbrdfIO", "reload_textures",
           "imageio", "materialgan", "scripts", "loss_mode", "upsample_textures", "synthetic", "LOSS_L2", "LOSS_LOG_L2", "SvbrdfHipError"]
