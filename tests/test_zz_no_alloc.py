"""afmm2d_direct_rows_device_noalloc (no allocation, static scratch) against
afmm2d_direct_rows on the same inputs: bit-identical, on two calls in a row
(the second with fewer rows), through the C ABI with device arrays."""
import ctypes

import numpy as np
import pytest

from conftest import uniform_points

pytestmark = pytest.mark.gpu


def _hip():
    hip = ctypes.CDLL("libamdhip64.so.7")
    vp = ctypes.c_void_p
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    hip.hipFree.argtypes = [vp]
    hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(vp), ctypes.c_uint]
    hip.hipStreamSynchronize.argtypes = [vp]
    hip.hipStreamDestroy.argtypes = [vp]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class Dev:
    """A device copy of a numpy array (freed on close)."""

    def __init__(self, hip, a):
        self.hip, self.nbytes = hip, a.nbytes
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(a.nbytes, 8)) == 0
        if a.nbytes:
            assert hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0

    def get(self, out):
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0
        return out

    def close(self):
        self.hip.hipFree(self.p)


@pytest.mark.parametrize("kernel,n", [("log", 20000), ("reg_inv_r", 7001), ("user", 5000)])
def test_noalloc_direct_rows_match_allocating_entry(engine, kernel, n):
    hip = _hip()
    xy, _ = uniform_points(n, 61)
    rng = np.random.default_rng(62)
    b = rng.uniform(-1, 1, n)
    A = engine.AFMM(n, 64, 10, xy, kernel=kernel, kparams=[0.0, 0.25, 0, 0] if kernel == "user" else None)
    db = Dev(hip, b)
    s = ctypes.c_void_p()
    assert hip.hipStreamCreateWithFlags(ctypes.byref(s), 1) == 0
    L = engine.lib()
    for nrows in (min(n, 3000), 500):  # the second call smaller than the first
        rows = np.sort(rng.choice(n, nrows, replace=False)).astype(np.int64)
        drows = Dev(hip, rows)
        out = np.full(nrows, np.nan)
        dout = Dev(hip, out)
        st = L.afmm2d_direct_rows_device_noalloc(A._h, nrows, drows.p, db.p, dout.p, s)
        assert st == 0, L.afmm2d_last_error()
        assert hip.hipStreamSynchronize(s) == 0
        got = dout.get(np.empty(nrows))
        ref = A.direct_rows(rows, b)
        np.testing.assert_array_equal(got, ref)
        drows.close()
        dout.close()
    # zero rows is a no-op; more rows than the scratch holds is refused
    assert L.afmm2d_direct_rows_device_noalloc(A._h, 0, None, db.p, None, None) == 0
    assert L.afmm2d_direct_rows_device_noalloc(A._h, (1 << 20) + 1, db.p, db.p, db.p, None) == \
        engine.ERR_UNSUPPORTED
    hip.hipStreamDestroy(s)
    db.close()
    A.close()
