"""The C ABI of libafmm2d.so on the CPU: exports, host-side kernel values and
argument validation (no GPU needed; nothing here launches a kernel)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "afmm2d.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(afmm2d_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_what_the_binding_lists(engine):
    assert declared_functions() == sorted(engine.EXPORTS)


def test_library_exports_every_declared_symbol(engine):
    lib = ctypes.CDLL(engine.library_path())
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_status_strings(engine):
    L = engine.lib()
    for s in range(8):
        assert L.afmm2d_status_string(s).decode() not in ("", "unknown status")
    assert L.afmm2d_status_string(99).decode() == "unknown status"


@pytest.mark.parametrize("kernel,kid,kp", [
    ("log", 0, None), ("inv_r", 1, None), ("gaussian", 2, [0.02, 0, 0, 0]),
    ("reg_inv_r", 3, None), ("user", 4, None), ("user", 4, [3.0, 0, 0, 0]),
    ("user", 4, [0.0, 0.25, 0, 0]), ("user_sym", 5, [0.0, 0.25, 0, 0]),
    ("user_anti", 6, [0.0, 0.25, 0, 0])])
def test_host_kernel_matches_oracle(engine, oracle_mod, kernel, kid, kp):
    rng = np.random.default_rng(kid)
    for _ in range(200):
        xi, yi, xj, yj = rng.uniform(-1, 1, 4)
        a = 0.25
        got = engine.kernel_value(kernel, xi, yi, xj, yj, a=a, kparams=kp)
        ref = oracle_mod.kernel_value(kid, xi, yi, xj, yj, a=a, kparams=kp)
        assert got == ref
    assert engine.kernel_value(kernel, 0.1, 0.2, 0.1, 0.2, a=0.25, kparams=kp) == \
        oracle_mod.kernel_value(kid, 0.1, 0.2, 0.1, 0.2, a=0.25, kparams=kp)


def _create(engine, n, leaf, tol_pow, loc, kid, layout=0):
    h = ctypes.c_void_p()
    kp = np.zeros(4)
    p = loc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if loc is not None else None
    st = engine.lib().afmm2d_create(n, leaf, tol_pow, p, layout, kid,
                                    kp.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), -1,
                                    ctypes.byref(h))
    return st, h


def test_create_rejects_bad_arguments(engine):
    loc = np.zeros((4096, 2))
    assert _create(engine, 4096, 64, 8, None, 0)[0] == engine.ERR_INVALID_ARGUMENT
    assert _create(engine, 4096, 64, 8, loc, 8)[0] == engine.ERR_INVALID_ARGUMENT
    assert _create(engine, 4096, 64, 8, loc, 0, layout=5)[0] == engine.ERR_INVALID_ARGUMENT
    assert _create(engine, 4096, 0, 8, loc, 0)[0] == engine.ERR_INVALID_ARGUMENT
    assert _create(engine, 4096, 64, -1, loc, 0)[0] == engine.ERR_INVALID_ARGUMENT
    # nLevels = log4(N/M) < 2: the reference has no interaction list there
    st, _ = _create(engine, 1000, 100, 8, loc, 0)
    assert st == engine.ERR_LEVELS
    assert "nLevels" in engine.lib().afmm2d_last_error().decode()
    st, _ = _create(engine, 100, 100, 8, loc, 0)
    assert st == engine.ERR_LEVELS


def test_null_handle_calls_fail_cleanly(engine):
    L = engine.lib()
    assert L.afmm2d_assemble(None) == engine.ERR_INVALID_ARGUMENT
    assert L.afmm2d_destroy(None) == engine.OK
    lv = ctypes.c_int()
    assert L.afmm2d_levels(None, ctypes.byref(lv)) == engine.ERR_INVALID_ARGUMENT


def test_levels_formula(engine, oracle_mod):
    for n, m in [(4096, 64), (1 << 20, 64), (6400, 200), (1000000, 100), (1 << 23, 128)]:
        assert engine.n_levels(n, m) == oracle_mod.nlevels(n, m)


def test_user_kernel_header_plugs_in(tmp_path):
    """The plug-in path a user takes: build the library with another kernel
    header (here a multiquadric, sqrt(R^2 + c^2), c = kparams[0]) and get that
    kernel as id 4 (host evaluation; the device kernels include the same
    header)."""
    from afmm2d_amd import build as B
    hdr = tmp_path / "multiquadric.h"
    hdr.write_text(
        "AFMM2D_USER_FN double afmm2d_user_kernel(double xi, double yi, double xj, double yj,\n"
        "                                         const double* kp, double a) {\n"
        "  const double c = kp[0] > 0.0 ? kp[0] : 0.1;\n"
        "  const double dx = xi - xj, dy = yi - yj;\n"
        "  return sqrt(dx * dx + dy * dy + c * c);\n"
        "}\n")
    lib_path = B.build(user_kernel=str(hdr), lib_path=str(tmp_path / "libafmm2d_mq.so"),
                       build_dir=str(tmp_path / "obj"), example=False)
    L = ctypes.CDLL(lib_path)
    L.afmm2d_kernel_value.restype = ctypes.c_int
    L.afmm2d_kernel_value.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                      ctypes.c_double] + [ctypes.c_double] * 4 + \
        [ctypes.POINTER(ctypes.c_double)]
    v = ctypes.c_double()
    assert L.afmm2d_kernel_value(4, None, 0.0, 0.1, 0.2, 0.4, -0.3, ctypes.byref(v)) == 0
    assert abs(v.value - np.sqrt(0.3 ** 2 + 0.5 ** 2 + 0.01)) <= 1e-15
    kp = (ctypes.c_double * 4)(0.5, 0.0, 0.0, 0.0)
    assert L.afmm2d_kernel_value(4, kp, 0.0, 0.1, 0.2, 0.4, -0.3, ctypes.byref(v)) == 0
    assert abs(v.value - np.sqrt(0.3 ** 2 + 0.5 ** 2 + 0.25)) <= 1e-15
    assert L.afmm2d_kernel_value(7, None, 0.0, 0.1, 0.2, 0.4, -0.3, ctypes.byref(v)) == 1
