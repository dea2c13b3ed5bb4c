"""Builds libafmm2d.so (HIP for gfx950) in-tree with hipcc.

Usage: python -m afmm2d_amd.build [--user-kernel HEADER]   (or afmm2d_amd.build.build())

The user kernel (kernel id 4, AFMM2D_KERNEL_USER) is compiled in from a
header defining afmm2d_user_kernel(); the default is
user_kernels/anisotropic_log.h, which also documents the contract.  The
environment variable AFMM2D_USER_KERNEL names another header.
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

SOURCES = ["api.cpp", "kd_sort.hip", "assemble.hip", "matvec.hip", "blocked.hip",
           "rank_model.hip"]
DEFAULT_USER_KERNEL = os.path.join(HERE, "user_kernels", "anisotropic_log.h")


def user_kernel_header(path=None):
    path = path or os.environ.get("AFMM2D_USER_KERNEL") or DEFAULT_USER_KERNEL
    path = os.path.abspath(path)
    if not os.path.isfile(path):
        raise RuntimeError("user kernel header not found: %s" % path)
    return path
ARCH = os.environ.get("AFMM2D_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def common_flags(user_kernel=None):
    # Packed FP32 instructions are disabled for every device pass.
    return [
        '-DAFMM2D_USER_KERNEL_HEADER="%s"' % user_kernel_header(user_kernel),
        "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics", "-Wno-unused-command-line-argument",
        "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    ]


def _compile(src, user_kernel=None, build_dir=BUILD):
    out = os.path.join(build_dir, os.path.basename(src) + ".o")
    lang = ["-x", "hip"]
    cmd = [hipcc()] + common_flags(user_kernel) + lang + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build(verbose=False, user_kernel=None, lib_path=LIB, build_dir=BUILD, example=True):
    """Compiles and links the library to lib_path (the in-tree library by
    default), with the user kernel from the header user_kernel."""
    os.makedirs(build_dir, exist_ok=True)
    user_kernel_header(user_kernel)
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda src: _compile(src, user_kernel, build_dir), SOURCES))
    tmp = lib_path + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, lib_path)
    if example and lib_path == LIB and os.path.exists(EXAMPLE_SRC):
        cmd = [hipcc(), "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), EXAMPLE_SRC,
               "-o", EXAMPLE_BIN, "-L" + HERE, "-lafmm2d", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("example build failed: %s\n%s" % (r.stdout, r.stderr))
    if verbose:
        print("built", lib_path)
    return lib_path


if __name__ == "__main__":
    uk = None
    if "--user-kernel" in sys.argv:
        uk = sys.argv[sys.argv.index("--user-kernel") + 1]
    build(verbose=True, user_kernel=uk)
    sys.exit(0)
