"""Builds libafmm2d.so (HIP for gfx950) in-tree with hipcc.

Usage: python -m afmm2d_amd.build   (or afmm2d_amd.build.build())
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libafmm2d.so")
EXAMPLE_SRC = os.path.join(HERE, "examples", "test_afmm2d.cpp")
EXAMPLE_BIN = os.path.join(HERE, "examples", "test_afmm2d")

SOURCES = ["api.cpp", "kd_sort.hip", "assemble.hip", "matvec.hip"]
ARCH = os.environ.get("AFMM2D_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def common_flags():
    # Packed FP32 instructions are disabled for every device pass.
    return [
        "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics", "-Wno-unused-command-line-argument",
        "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    ]


def _compile(src):
    out = os.path.join(BUILD, os.path.basename(src) + ".o")
    lang = ["-x", "hip"]
    cmd = [hipcc()] + common_flags() + lang + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if os.path.exists(EXAMPLE_SRC):
        cmd = [hipcc(), "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), EXAMPLE_SRC,
               "-o", EXAMPLE_BIN, "-L" + HERE, "-lafmm2d", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("example build failed: %s\n%s" % (r.stdout, r.stderr))
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
