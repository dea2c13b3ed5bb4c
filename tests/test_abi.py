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


@pytest.mark.parametrize("kernel,kid", [("log", 0), ("inv_r", 1), ("gaussian", 2), ("reg_inv_r", 3)])
def test_host_kernel_matches_oracle(engine, oracle_mod, kernel, kid):
    rng = np.random.default_rng(kid)
    kp = [0.02, 0, 0, 0] if kernel == "gaussian" else None
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
    assert _create(engine, 4096, 64, 8, loc, 7)[0] == engine.ERR_INVALID_ARGUMENT
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
