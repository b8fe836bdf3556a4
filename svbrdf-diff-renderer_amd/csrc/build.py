"""Builds libsvbrdf_hip.so (gfx950) next to the package.

    python svbrdf-diff-renderer_amd/csrc/build.py [--asm | --sanitize]

hipcc cross-compiles without a GPU.  ``--asm`` also writes the gfx950
assembly of the device code to csrc/_build/ for instruction counting;
``--sanitize`` builds the host side with ASan + UBSan instead (csrc/_build/).
"""
import os
import pathlib
import shutil
import subprocess
import sys

CSRC = pathlib.Path(__file__).resolve().parent
PKG = CSRC.parent
ROOT = PKG.parent
LIB = PKG / "libsvbrdf_hip.so"
SOURCES = [CSRC / "svbrdf_hip.hip", CSRC / "codec.hip", CSRC / "conv.hip", CSRC / "common.hip"]
HEADERS = [CSRC / "brdf.hpp", CSRC / "common.hpp", ROOT / "include" / "svbrdf_hip.h"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off: some of their forms have given wrong
    # results on MI355X now and then, with no error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]


def build_sanitized(verbose=False):
    """The library with AddressSanitizer and UndefinedBehaviorSanitizer on
    its host code (each -fsanitize= behind -Xarch_host: the device code is
    built as usual, without them), into csrc/_build/libsvbrdf_hip_san.so.
    Load it with SVBRDF_AMD_LIB pointing at it and clang's ASan runtime
    preloaded into the Python process."""
    out = CSRC / "_build"
    out.mkdir(exist_ok=True)
    lib = out / "libsvbrdf_hip_san.so"
    tmp = lib.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-shared", "-shared-libsan",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
           "-Wno-unused-command-line-argument", f"-I{ROOT / 'include'}", *map(str, SOURCES), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    return lib


def asan_runtime():
    """Path of clang's shared ASan runtime that matches hipcc."""
    r = subprocess.run([HIPCC, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path):
        return path
    import glob
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def build(asm=False, verbose=False):
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, *map(str, SOURCES), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if asm:
        out = CSRC / "_build"
        out.mkdir(exist_ok=True)
        subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only", "-S",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        f"-I{ROOT / 'include'}", str(SOURCES[0]), "-o", str(out / "svbrdf_hip.s")],
                       check=True)
        subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only", "-S",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        f"-I{ROOT / 'include'}", str(SOURCES[1]), "-o", str(out / "codec.s")],
                       check=True)
    return LIB


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        print("built", build_sanitized(verbose=True))
    else:
        build(asm="--asm" in sys.argv, verbose=True)
        print("built", LIB)
