"""C-ABI checks that need no GPU: the library loads, exports every entry point
include/svbrdf_hip.h declares, and rejects bad arguments on the host (before
any HIP call) with the documented status codes."""
import ctypes
import re

import pytest

from tests.conftest import ROOT

HEADER = ROOT / "include" / "svbrdf_hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(svbrdf_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib(pkg):
    return pkg._lib.load()


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for n in ("svbrdf_render", "svbrdf_render_backward", "svbrdf_fused_step", "svbrdf_prepare_targets",
              "svbrdf_step_state_init", "svbrdf_abi_version", "svbrdf_last_error",
              "svbrdf_fused_step_flush"):
        assert n in names


def test_library_exports_every_declared_symbol(lib, pkg):
    for n in declared_functions():
        assert hasattr(lib, n), n
        assert n in pkg._lib.SIGNATURES, f"{n} has no ctypes signature"


def test_queries(lib, pkg):
    assert lib.svbrdf_abi_version() == 12 == pkg._lib.ABI_VERSION
    # the host-side view of svbrdf_step_state matches the C struct
    assert lib.svbrdf_step_state_bytes() == pkg.optim._STATE_DTYPE.itemsize == 96
    assert lib.svbrdf_max_res() >= 4096
    assert lib.svbrdf_max_views() >= 16
    # 65 arrival counters on 128-byte lines (8448 bytes) + two buffers of one
    # float4 partial per 256-pixel block, each rounded up to 256 bytes
    assert lib.svbrdf_fused_step_workspace_bytes(1024) == 8448 + 2 * (1024 * 1024 // 256 * 16)
    assert lib.svbrdf_fused_step_workspace_bytes(1) == 8448 + 2 * 256
    assert lib.svbrdf_fused_step_workspace_bytes(0) == 0
    # one float4 light-power partial per 64-pixel block of the adjoint
    assert lib.svbrdf_render_backward_workspace_bytes(13) == 3 * 16
    assert lib.svbrdf_render_backward_workspace_bytes(1024) == 1024 * 1024 // 64 * 16
    # padded plane stride: res^2 rounded up to 64 floats, plus 448
    assert lib.svbrdf_plane_stride(1024) == 1024 * 1024 + 448
    assert lib.svbrdf_plane_stride(1023) == 1046592 + 448 and 1046592 >= 1023 * 1023
    assert lib.svbrdf_plane_stride(1) == 64 + 448
    assert lib.svbrdf_plane_stride(0) == 0 and lib.svbrdf_plane_stride(lib.svbrdf_max_res() + 1) == 0
    # 9 planes at the largest resolution stay inside one 2 GiB buffer range
    assert 9 * 4 * lib.svbrdf_plane_stride(lib.svbrdf_max_res()) < 2 ** 31


def _fake(n=1):
    # never dereferenced: validation fails before any launch
    return ctypes.c_void_p(0x1000 * n)


@pytest.mark.parametrize("res,views,code", [(0, 9, -2), (-4, 9, -2), (1 << 20, 9, -2), (8, 0, -2), (8, 100000, -2)])
def test_render_rejects_bad_shapes(lib, res, views, code):
    rc = lib.svbrdf_render(_fake(1), res, views, ctypes.c_float(17.0), _fake(2), _fake(3), _fake(4), _fake(5), None)
    assert rc == code
    assert b"svbrdf_render" in lib.svbrdf_last_error()


def test_render_rejects_null_pointers(lib):
    assert lib.svbrdf_render(None, 8, 9, ctypes.c_float(17.0), _fake(2), _fake(3), _fake(4), _fake(5), None) == -1
    assert lib.svbrdf_render(_fake(1), 8, 9, ctypes.c_float(17.0), _fake(2), _fake(3), _fake(4), None, None) == -1


def test_backward_requires_workspace_for_light_gradient(lib):
    rc = lib.svbrdf_render_backward(_fake(1), 64, 9, ctypes.c_float(17.0), _fake(2), _fake(3), _fake(4), _fake(5),
                                    _fake(6), _fake(7), None, 0, None)
    assert rc == -3
    assert b"workspace" in lib.svbrdf_last_error()
    # one float4 partial per 64-pixel block, 16-byte aligned
    need = lib.svbrdf_render_backward_workspace_bytes(64)
    assert need == 64 * 64 // 64 * 16
    rc = lib.svbrdf_render_backward(_fake(1), 64, 9, ctypes.c_float(17.0), _fake(2), _fake(3), _fake(4), _fake(5),
                                    _fake(6), _fake(7), ctypes.c_void_p(0x10008), need, None)
    assert rc == -1
    assert b"16-byte aligned" in lib.svbrdf_last_error()


def _step(lib, **kw):
    a = dict(tex=_fake(1), m=_fake(2), v=_fake(3), tgt=_fake(4), res=64, tps=0, ps=0, views=9, cam=_fake(5),
             light=_fake(6),
             pow_=_fake(7), mode=1, light_opt=0, lr=0.01, b1=0.9, b2=0.999, eps=1e-8, state=_fake(8),
             hist=None, hlen=0, ws=_fake(9), wsb=lib.svbrdf_fused_step_workspace_bytes(64))
    a.update(kw)
    return lib.svbrdf_fused_step(a["tex"], a["m"], a["v"], a["tgt"], a["res"], a["tps"], a["ps"], a["views"],
                                 ctypes.c_float(17.0),
                                 a["cam"], a["light"], a["pow_"], a["mode"], a["light_opt"],
                                 ctypes.c_double(a["lr"]), ctypes.c_double(a["b1"]), ctypes.c_double(a["b2"]),
                                 ctypes.c_double(a["eps"]), a["state"], a["hist"], a["hlen"], a["ws"], a["wsb"], None)


def test_fused_step_validation(lib):
    assert _step(lib, mode=7) == -1
    assert _step(lib, b1=1.0) == -1
    assert _step(lib, lr=float("nan")) == -1
    assert _step(lib, state=None) == -1
    assert _step(lib, wsb=16) == -3
    assert _step(lib, ws=ctypes.c_void_p(0x1004)) == -1      # misaligned workspace
    assert _step(lib, ws=ctypes.c_void_p(0x1080)) == -1      # only 128-byte aligned
    assert _step(lib, res=0) == -2
    assert _step(lib, ps=64 * 64 - 1) == -1                  # planes overlap
    assert _step(lib, tps=100) == -1
    assert _step(lib, ps=60_000_000) == -2                   # 9 planes beyond 2 GiB
    assert b"plane_stride" in lib.svbrdf_last_error()


def test_flush_validation(lib):
    need = lib.svbrdf_fused_step_workspace_bytes(64)
    assert lib.svbrdf_fused_step_flush(None, 64, 9, None, 0, _fake(9), need, None) == -1
    assert lib.svbrdf_fused_step_flush(_fake(8), 64, 9, None, 0, None, need, None) == -1
    assert lib.svbrdf_fused_step_flush(_fake(8), 0, 9, None, 0, _fake(9), need, None) == -2
    assert lib.svbrdf_fused_step_flush(_fake(8), 64, 0, None, 0, _fake(9), need, None) == -2
    assert lib.svbrdf_fused_step_flush(_fake(8), 64, 9, None, 0, _fake(9), need - 1, None) == -3
    assert b"svbrdf_fused_step_flush" in lib.svbrdf_last_error()


def test_prepare_and_state_validation(lib):
    assert lib.svbrdf_prepare_targets(None, 10, 1, _fake(1), None) == -1
    assert lib.svbrdf_prepare_targets(_fake(1), 10, 5, _fake(1), None) == -1
    assert lib.svbrdf_prepare_targets(_fake(1), 0, 1, _fake(1), None) == 0     # nothing to do
    assert lib.svbrdf_step_state_init(None, 0.9, 0.999, None) == -1
    assert lib.svbrdf_step_state_init(_fake(1), 0.9, 1.5, None) == -1


def test_python_wrapper_refuses_cpu(pkg):
    import torch
    cl = [torch.zeros(9, 3), torch.zeros(9, 3), torch.ones(3)]
    with pytest.raises(RuntimeError, match="GPU only"):
        pkg.Microfacet(8, 9, 17.0, cl, "cpu")


def test_python_wrapper_keeps_reference_asserts(pkg):
    """Microfacet.eval asserts square, res-matching textures
    (src/microfacet.py:85-86) before any device work."""
    import torch
    r = pkg.Microfacet.__new__(pkg.Microfacet)
    r.res = 8
    with pytest.raises(AssertionError):
        r.eval(torch.zeros(1, 9, 8, 6))
    with pytest.raises(AssertionError):
        r.eval(torch.zeros(1, 9, 6, 6))


def test_sizes_beyond_the_limits_are_refused(lib):
    big = lib.svbrdf_max_res() + 1
    assert lib.svbrdf_max_res() == 7680          # 9 * 7680^2 * 4 bytes < 2^31
    assert lib.svbrdf_fused_step_workspace_bytes(big) == 0
    assert lib.svbrdf_render_backward_workspace_bytes(big) == 0
    assert lib.svbrdf_render(_fake(1), big, 1, ctypes.c_float(17.0), _fake(2), _fake(3), _fake(4), _fake(5),
                             None) == -2
    assert lib.svbrdf_render(_fake(1), 8, lib.svbrdf_max_views() + 1, ctypes.c_float(17.0), _fake(2), _fake(3),
                             _fake(4), _fake(5), None) == -2
    assert _step(lib, res=big) == -2


def test_noalloc_adjoint_refuses_sizes_beyond_its_scratch(lib):
    """svbrdf_render_backward_noalloc serves res <= 4096 (its 4 MiB array);
    larger sizes are refused on the host before any HIP call."""
    f = _fake()
    assert lib.svbrdf_render_backward_noalloc(f, 4097, 9, 17.0, f, f, f, f, f, f, None) == -2
    assert b"4096" in lib.svbrdf_last_error()


def test_conv_workspace_plan(lib):
    """svbrdf_conv3x3 splits the channel sum only where the output tiles
    alone leave the device short of blocks: 512 -> 512 channels at 4 x 4 runs
    in slices of the input channels whose partial sums are [slices][1, 512,
    4, 4]; a 256^2 layer of 9 images and the direct (<= 8 output channels)
    kernel need none; a short workspace is refused before any launch."""
    plane = 512 * 16 * 4
    need = lib.svbrdf_conv3x3_workspace_bytes(1, 512, 4, 4, 512)
    assert need % plane == 0 and 2 <= need // plane <= 64
    assert lib.svbrdf_conv3x3_workspace_bytes(9, 64, 256, 256, 64) == 0
    assert lib.svbrdf_conv3x3_workspace_bytes(9, 64, 256, 256, 3) == 0
    assert lib.svbrdf_conv3x3_workspace_bytes(0, 64, 8, 8, 64) == 0
    f = _fake()
    assert lib.svbrdf_conv3x3(f, 1, 512, 4, 4, f, 512, None, None, 0.0, 0, f, None, 0, None) == -3
    assert b"workspace" in lib.svbrdf_last_error()
    assert lib.svbrdf_conv3x3(f, 1, 512, 4, 4, f, 512, None, None, 0.0, 0, f, f, need - 4, None) == -3
    assert lib.svbrdf_conv3x3(f, 1, 0, 4, 4, f, 512, None, None, 0.0, 0, f, None, 0, None) == -2
    assert lib.svbrdf_conv3x3(f, 1, 8, 4, 4, f, 8, None, None, 0.0, 7, f, None, 0, None) == -1


def test_conv_weight_gradient_plan_and_validation(lib):
    """svbrdf_conv3x3_wgrad splits the pixel sum only where the channel tiles
    are few and the pixels many: 128 -> 128 channels at 256^2 runs in slices
    whose partial sums are [slices][128][128][9]; 512 -> 512 at 4 x 4 (one
    16-pixel chunk) needs no workspace.  Bad shapes and a short workspace are
    refused before any launch; so are bad pooling and gated-gradient calls."""
    per = 128 * 128 * 9 * 4
    need = lib.svbrdf_conv3x3_wgrad_workspace_bytes(1, 128, 256, 256, 128)
    assert need % per == 0 and 2 <= need // per <= 256
    assert lib.svbrdf_conv3x3_wgrad_workspace_bytes(1, 512, 4, 4, 512) == 0
    assert lib.svbrdf_conv3x3_wgrad_workspace_bytes(1, 0, 4, 4, 512) == 0
    f = _fake()
    assert lib.svbrdf_conv3x3_wgrad(f, f, 1, 128, 256, 256, 128, f, None, 0, None) == -3
    assert lib.svbrdf_conv3x3_wgrad(f, f, 1, 128, 256, 256, 128, f, f, need - 4, None) == -3
    assert lib.svbrdf_conv3x3_wgrad(f, f, 0, 8, 4, 4, 8, f, None, 0, None) == -2
    assert lib.svbrdf_conv3x3_wgrad(None, f, 1, 8, 4, 4, 8, f, None, 0, None) == -1
    assert lib.svbrdf_avgpool2(f, 1, 1, 4, f, None) == -2
    assert lib.svbrdf_avgpool2_adjoint(None, 1, 4, 4, f, None) == -1
    assert lib.svbrdf_conv3x3_input_grad(f, f, 9, 1, 8, 4, 4, f, 8, f, None, 0, None) == -1   # activation
    assert lib.svbrdf_conv3x3_input_grad(f, None, 1, 1, 8, 4, 4, f, 8, f, None, 0, None) == -1


def test_modulation_grad_validation(lib):
    """svbrdf_conv3x3_modulation_grad refuses null pointers, bad channel counts
    and a workspace shorter than 2 cout floats (or misaligned) before any
    launch."""
    f = _fake()
    assert lib.svbrdf_conv3x3_modulation_grad(None, 8, 8, f, 1.0, 1, f, f, f, 64, None) == -1
    assert lib.svbrdf_conv3x3_modulation_grad(f, 8, 8, None, 1.0, 1, f, f, f, 64, None) == -1
    assert lib.svbrdf_conv3x3_modulation_grad(f, 0, 8, f, 1.0, 1, f, f, f, 64, None) == -2
    assert lib.svbrdf_conv3x3_modulation_grad(f, 8, 8, f, 1.0, 1, f, f, f, 63, None) == -3
    assert lib.svbrdf_conv3x3_modulation_grad(f, 8, 8, f, 1.0, 1, f, f, ctypes.c_void_p(0x10004), 64, None) == -1
    assert b"aligned" in lib.svbrdf_last_error()


def test_act_grad_validation(lib):
    f = _fake()
    assert lib.svbrdf_act_grad(None, f, 16, 1.0, 0.0, f, None) == -1
    assert lib.svbrdf_act_grad(f, f, 0, 1.0, 0.0, f, None) == 0      # nothing to do
