"""The library's host code (argument validation, the batch planner, the
error state) under AddressSanitizer and UndefinedBehaviorSanitizer: the
host-side tests (tests/test_abi.py, tests/test_host.py) run in a child
process against the sanitizer build of csrc/build.py --sanitize, with
clang's ASan runtime preloaded.  Any sanitizer report fails the child."""
import importlib.util
import os
import subprocess
import sys

import pytest

from tests.conftest import ROOT

BUILD = ROOT / "svbrdf-diff-renderer_amd" / "csrc" / "build.py"


def _build_module():
    spec = importlib.util.spec_from_file_location("_svbrdf_build_san", BUILD)
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    return b


@pytest.mark.timeout(900)
def test_host_code_under_asan_and_ubsan(tmp_path):
    b = _build_module()
    runtime = b.asan_runtime()
    if runtime is None:
        pytest.skip("clang's ASan runtime is not installed")
    lib = b.CSRC / "_build" / "libsvbrdf_hip_san.so"
    sources = [*b.SOURCES, *b.HEADERS]
    if not lib.exists() or lib.stat().st_mtime < max(p.stat().st_mtime for p in sources):
        b.build_sanitized()
    syms = subprocess.run(["nm", "-D", str(lib)], capture_output=True, text=True).stdout
    assert "__asan_" in syms and "__ubsan_" in syms, "the sanitizer build is not instrumented"
    env = dict(os.environ, SVBRDF_AMD_LIB=str(lib), LD_PRELOAD=runtime,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        str(ROOT / "tests" / "test_abi.py"), str(ROOT / "tests" / "test_host.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=800)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert " passed" in out
