"""ctypes binding of libsvbrdf_hip.so (the C ABI in include/svbrdf_hip.h).

The library is the only compute path: if it is missing or a call fails, the
wrapper raises.  There is no CPU or eager-torch fallback.
"""
import ctypes
import os
import pathlib

import torch

_PKG = pathlib.Path(__file__).resolve().parent
# SVBRDF_AMD_LIB selects another build of the same library (the host-side
# sanitizer build of csrc/build.py --sanitize, for the CPU tests)
LIB_PATH = pathlib.Path(os.environ.get("SVBRDF_AMD_LIB") or _PKG / "libsvbrdf_hip.so")

# the SVBRDF_ABI_VERSION these signatures describe (include/svbrdf_hip.h)
ABI_VERSION = 12

OK = 0
EINVAL = -1
ESHAPE = -2
ESPACE = -3
EHIP = -4

LOSS_L2 = 0
LOSS_LOG_L2 = 1
IMREAD_NONE, IMREAD_SRGB, IMREAD_ROUGH, IMREAD_NORMAL = 0, 1, 2, 3
TERMS_EVAL, NDF_BECKMANN, FRESNEL_SCHLICK = 0, 1, 2
ACT_NONE, ACT_RELU, ACT_LRELU = 0, 1, 2

_c_float_p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p

# name -> (restype, argtypes): every symbol declared in include/svbrdf_hip.h
SIGNATURES = {
    "svbrdf_abi_version": (ctypes.c_int, []),
    "svbrdf_last_error": (ctypes.c_char_p, []),
    "svbrdf_step_state_bytes": (ctypes.c_size_t, []),
    "svbrdf_max_res": (ctypes.c_int, []),
    "svbrdf_max_views": (ctypes.c_int, []),
    "svbrdf_render": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, _vp, _vp, _vp, _vp, _vp]),
    "svbrdf_render_backward_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int]),
    "svbrdf_render_backward": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, _vp, _vp, _vp,
                                              _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "svbrdf_render_terms": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, _vp, _vp, _vp,
                                           ctypes.c_int, _vp, _vp]),
    "svbrdf_render_backward_terms": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, _vp, _vp, _vp,
                                                    ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "svbrdf_render_backward_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, _vp, _vp,
                                                      _vp, _vp, _vp, _vp, _vp]),
    "svbrdf_prepare_targets": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_int, _vp, _vp]),
    "svbrdf_step_state_init": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _vp]),
    "svbrdf_fused_step_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int]),
    "svbrdf_plane_stride": (ctypes.c_size_t, [ctypes.c_int]),
    "svbrdf_fused_step": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t,
                                         ctypes.c_int, ctypes.c_float, _vp, _vp,
                                         _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, _vp, _vp, ctypes.c_int, _vp,
                                         ctypes.c_size_t, _vp]),
    # batched step (svbrdf_material tables are passed as pointers)
    "svbrdf_fused_step_flush": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp,
                                               ctypes.c_size_t, _vp]),
    "svbrdf_resize_lanczos4_u8": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int,
                                                 ctypes.c_int, _vp]),
    "svbrdf_upsample2x_bilinear": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_size_t, _vp, ctypes.c_size_t,
                                                  ctypes.c_int, _vp]),
    "svbrdf_images_from_u8": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, _vp, _vp]),
    "svbrdf_textures_from_u8": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "svbrdf_textures_to_u8": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]),
    "svbrdf_images_to_u8": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, _vp, _vp]),
    "svbrdf_batch_tile_count": (ctypes.c_size_t, [_vp, ctypes.c_int, ctypes.c_int]),
    "svbrdf_batch_plan": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_size_t]),
    "svbrdf_batch_workspace_bytes": (ctypes.c_size_t, [ctypes.c_size_t]),
    "svbrdf_batch_max_blocks_per_cu": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "svbrdf_batch_step": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double, _vp, _vp, _vp,
                                         ctypes.c_int, _vp, ctypes.c_size_t, _vp]),
    # convolutions of the MaterialGAN caller's networks
    "svbrdf_conv3x3_pack": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_float, ctypes.c_int,
                                           ctypes.c_int, _vp, _vp]),
    "svbrdf_act_grad": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, ctypes.c_float, ctypes.c_float, _vp, _vp]),
    "svbrdf_conv3x3_modulation_grad": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_float,
                                                      ctypes.c_int, _vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "svbrdf_conv3x3_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int] * 5),
    "svbrdf_conv3x3": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                                      ctypes.c_int, _vp, _vp, ctypes.c_float, ctypes.c_int, _vp, _vp,
                                      ctypes.c_size_t, _vp]),
    "svbrdf_conv3x3_input_grad": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, ctypes.c_size_t, _vp]),
    "svbrdf_conv3x3_wgrad_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int] * 5),
    "svbrdf_conv3x3_wgrad": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, _vp, _vp, ctypes.c_size_t, _vp]),
    "svbrdf_avgpool2": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "svbrdf_avgpool2_adjoint": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "svbrdf_upsample2_blur": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "svbrdf_upsample2_blur_adjoint": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]),
}

_lib = None


class SvbrdfHipError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def load():
    """Loads the shared library once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with `python __graft_entry__.py build` "
                          "(or svbrdf-diff-renderer_amd/csrc/build.py)")
    lib = ctypes.CDLL(os.fspath(LIB_PATH))
    lib.svbrdf_abi_version.restype = ctypes.c_int
    version = lib.svbrdf_abi_version()
    if version != ABI_VERSION:
        # argument lists differ between versions: calling through the wrong
        # signatures would pass garbage pointers to the kernels
        raise ImportError(f"{LIB_PATH} implements ABI {version}, these bindings need {ABI_VERSION}: rebuild it "
                          "with `python __graft_entry__.py build`")
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def call(name, *args):
    """Calls an int-returning entry point and raises on a non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != OK:
        msg = lib.svbrdf_last_error().decode(errors="replace")
        if rc in (EINVAL, ESHAPE, ESPACE):
            raise ValueError(f"{name}: {msg}")
        raise SvbrdfHipError(name, rc, msg)


def query(name, *args):
    return getattr(load(), name)(*args)


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def cuda_device(device, what="svbrdf_amd"):
    """The torch.device of a GPU with its index filled in: "cuda" becomes
    "cuda:<current device>", so that it compares equal to the device of the
    tensors created on it.  Raises RuntimeError for a non-GPU device (there is
    no CPU path)."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise RuntimeError(f"{what} on {dev}: the HIP library runs on a GPU only (there is no CPU path)")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


def stream_handle(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
