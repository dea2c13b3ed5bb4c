"""The blocked matvec (several right-hand sides per pass on the matrix cores,
FP64 / FP32 / BF16 factors) against the single-RHS FP64 passes, the CPU
oracle and the FP64 direct sum, on the same inputs.

Bars: FP64 columns within 1e-12 relative of the single-RHS result (the same
factors summed in another order; observed ~1e-15) and 1e-11 of the oracle.
FP32 and BF16 apply an orthonormal basis of the same operator rounded to the
precision; their error against the FP64 operator is bounded per case (FP32
a few 1e-7, BF16 about 1e-3 .. 1e-2), and against the direct sum it is that
plus the FP64 operator's own error."""
import ctypes

import numpy as np
import pytest

from conftest import uniform_points

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def col_rel(Y, Yref):
    return max(rel(Y[:, c], Yref[:, c]) for c in range(Y.shape[1]))


def print_case(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("n,leaf,tol_pow,kernel,nrhs", [
    (4096, 64, 10, "log", 1),
    (4096, 64, 10, "log", 16),
    (20000, 64, 12, "log", 5),
    (20000, 64, 12, "log", 32),
    (30000, 64, 10, "inv_r", 17),
    (12345, 48, 11, "gaussian", 8),
    (50000, 100, 12, "reg_inv_r", 3),   # ranks above 64: several row groups per box
    (5000, 64, 10, "user", 2),
])
def test_fp64_blocked_matches_single_rhs_and_oracle(engine, oracle_mod, n, leaf, tol_pow, kernel,
                                                    nrhs):
    print_case("blocked fp64", n, leaf, tol_pow, kernel, nrhs)
    xy, _ = uniform_points(n, n + nrhs)
    rng = np.random.default_rng(n)
    Q = rng.uniform(-1, 1, size=(n, nrhs))
    A = engine.AFMM(n, leaf, tol_pow, xy, kernel=kernel)
    A.assemble()
    Yb = A.matvec(Q, "fp64", blocked=True)
    Ys = A.matvec(Q, "fp64")  # one column at a time, single-RHS passes
    assert col_rel(Yb, Ys) <= 1e-12
    kid = engine.KERNELS[kernel]
    T = oracle_mod.Tree(xy, leaf, tol_pow, kid)
    T.assemble()
    for c in range(min(nrhs, 3)):
        assert rel(Yb[:, c], T.matvec(Q[:, c])) <= 1e-11
    st = A.stats()
    # FP64: every block once, padded to 4 x 4 tiles; FP32/BF16: both directions
    assert st["m2l_bytes"] // 8 <= st["blocked_m2l_coeffs"] <= 1.5 * st["m2l_bytes"] // 8
    assert st["blocked_m2l_coeffs_bf16"] == 2 * st["m2l_bytes"] // 8
    A.close()
    T.close()


def test_fp64_blocked_deterministic_and_linear(engine):
    n = 20000
    xy, _ = uniform_points(n, 11)
    rng = np.random.default_rng(12)
    Q = rng.standard_normal((n, 12))
    A = engine.AFMM(n, 64, 10, xy)
    A.assemble()
    Y1 = A.matvec(Q, "fp64", blocked=True)
    Y2 = A.matvec(Q, "fp64", blocked=True)
    np.testing.assert_array_equal(Y1, Y2)
    # a column does not depend on the others in its pass
    Y3 = A.matvec(np.asfortranarray(Q[:, [4, 0]]), "fp64", blocked=True)
    np.testing.assert_array_equal(Y3[:, 0], Y1[:, 4])
    np.testing.assert_array_equal(Y3[:, 1], Y1[:, 0])
    # linearity to rounding (amplified by the interpolative basis' large
    # coefficients, as in the single-RHS passes)
    yl = A.matvec(2.0 * Q[:, 0] - Q[:, 1], "fp64", blocked=True)
    assert rel(yl, 2.0 * Y1[:, 0] - Y1[:, 1]) <= 1e-12
    A.close()


def test_more_than_32_columns_and_leading_dimensions(engine):
    n = 8000
    xy, _ = uniform_points(n, 21)
    rng = np.random.default_rng(22)
    Q = rng.uniform(-1, 1, size=(n, 40))
    A = engine.AFMM(n, 64, 10, xy)
    A.assemble()
    Y = A.matvec(Q, "fp64", blocked=True)       # chunks of 32 + 8
    Ys = A.matvec(Q, "fp64")
    assert col_rel(Y, Ys) <= 1e-12
    # host arrays with ldb, ldy > N through the C ABI
    ld = n + 37
    B = np.zeros((ld, 5), order="F")
    B[:n] = Q[:, :5]
    Yp = np.full((ld + 3, 5), 7.0, order="F")
    dp = ctypes.POINTER(ctypes.c_double)
    st = engine.lib().afmm2d_matvec_ex(A._h, B.ctypes.data_as(dp), ld,
                                       Yp.ctypes.data_as(dp), ld + 3, 5,
                                       engine.MATVEC_BLOCKED)
    assert st == 0
    assert col_rel(Yp[:n], Ys[:, :5]) <= 1e-12
    assert (Yp[n:] == 7.0).all()  # padding rows untouched
    # invalid modes and leading dimensions are refused
    L = engine.lib()
    assert L.afmm2d_matvec_ex(A._h, B.ctypes.data_as(dp), n - 1, Yp.ctypes.data_as(dp), ld, 5, 0) \
        == engine.ERR_INVALID_ARGUMENT
    assert L.afmm2d_matvec_ex(A._h, B.ctypes.data_as(dp), ld, Yp.ctypes.data_as(dp), ld, 5, 3) \
        == engine.ERR_INVALID_ARGUMENT
    A.close()


@pytest.mark.parametrize("n,tol_pow,kernel,bar32,bar16", [
    (4096, 10, "log", 1e-6, 3e-2),
    (65536, 12, "log", 1e-6, 3e-2),
    (30000, 10, "inv_r", 1e-6, 3e-2),
    (20000, 10, "reg_inv_r", 1e-6, 3e-2),
    (12000, 10, "gaussian", 1e-6, 3e-2),
])
def test_reduced_precision_error(engine, n, tol_pow, kernel, bar32, bar16):
    """FP32 and BF16 against the FP64 operator (same tree and skeletons) and
    the FP64 direct sum; the errors are printed for the record."""
    print_case("reduced precision", n, tol_pow, kernel)
    xy, _ = uniform_points(n, 31 + n)
    rng = np.random.default_rng(32)
    Q = rng.uniform(-1, 1, size=(n, 4))
    A = engine.AFMM(n, 64, tol_pow, xy, kernel=kernel)
    A.assemble()
    Y64 = A.matvec(Q, "fp64", blocked=True)
    rows = np.arange(0, n, max(1, n // 2048))
    yd = A.direct_rows(rows, Q[:, 0])
    e64 = rel(Y64[rows, 0], yd)
    out = {}
    for prec, bar in (("fp32", bar32), ("bf16", bar16)):
        Y = A.matvec(Q, prec)
        d = col_rel(Y, Y64)
        e = rel(Y[rows, 0], yd)
        out[prec] = (d, e)
        assert d <= bar, (prec, d)
        assert e <= bar + 2 * e64, (prec, e, e64)
    print_case("  fp64 err %.2e | fp32 vs fp64 %.2e err %.2e | bf16 vs fp64 %.2e err %.2e" %
               (e64, out["fp32"][0], out["fp32"][1], out["bf16"][0], out["bf16"][1]))
    # FP32 is much closer to FP64 than BF16
    assert out["fp32"][0] * 100 < out["bf16"][0]
    A.close()


def test_reduced_precision_single_column_and_repeat(engine):
    n = 16384
    xy, q = uniform_points(n, 41)
    A = engine.AFMM(n, 64, 12, xy)
    A.assemble()
    A.prepare_precision("fp32")
    y1 = A.matvec(q, "fp32")
    y2 = A.matvec(q, "fp32")
    np.testing.assert_array_equal(y1, y2)
    y64 = A.computeMatVecProduct(q)
    assert rel(y1, y64) <= 1e-6
    A.close()


def test_rank_zero_far_field_blocked(engine):
    # every far-field rank is 0: the blocked passes reduce to the near field
    xy, _ = uniform_points(4096, 3)
    rng = np.random.default_rng(4)
    Q = rng.uniform(-1, 1, size=(4096, 3))
    A = engine.AFMM(4096, 64, 10, xy, kernel="gaussian", kparams=[1e-5, 0, 0, 0])
    A.assemble()
    Y = A.matvec(Q, "fp64", blocked=True)
    for c in range(3):
        assert rel(Y[:, c], A.direct(Q[:, c])) <= 1e-13
    Y32 = A.matvec(Q, "fp32")
    assert col_rel(Y32, Y) <= 1e-6
    A.close()


def test_nonsymmetric_user_kernel_blocked(engine):
    kp = [0.0, 0.25, 0.0, 0.0]
    n = 12000
    xy, _ = uniform_points(n, 99)
    rng = np.random.default_rng(98)
    Q = rng.uniform(-1, 1, size=(n, 6))
    A = engine.AFMM(n, 64, 10, xy, kernel="user", kparams=kp)
    A.assemble()
    Yb = A.matvec(Q, "fp64", blocked=True)
    Ys = A.matvec(Q, "fp64")
    assert col_rel(Yb, Ys) <= 1e-12
    Y32 = A.matvec(Q, "fp32")
    assert col_rel(Y32, Ys) <= 1e-6
    A.close()


def _hip():
    """The HIP runtime libafmm2d.so links (the same instance in this
    process), for device buffers and a stream of the test's own."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    vp = ctypes.c_void_p
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    hip.hipFree.argtypes = [vp]
    hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(vp), ctypes.c_uint]
    hip.hipStreamSynchronize.argtypes = [vp]
    hip.hipStreamDestroy.argtypes = [vp]
    return hip


def test_device_entry_on_a_stream(engine):
    """afmm2d_matvec_device_ex on device arrays with ldb, ldy > N on a
    non-blocking stream, then a host call on the null stream right after
    (ordered after the device call by the handle)."""
    n = 20000
    xy, _ = uniform_points(n, 51)
    rng = np.random.default_rng(52)
    Q = rng.uniform(-1, 1, size=(n, 16))
    A = engine.AFMM(n, 64, 10, xy)
    A.assemble()
    hip = _hip()
    ld = n + 5
    Bh = np.zeros((ld, 16), order="F")
    Bh[:n] = Q
    dB, dY, s = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dB), Bh.nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(dY), Bh.nbytes) == 0
    assert hip.hipStreamCreateWithFlags(ctypes.byref(s), 1) == 0  # non-blocking
    assert hip.hipMemcpy(dB, Bh.ctypes.data, Bh.nbytes, 1) == 0
    A.matvec_device_ex(dB.value, ld, dY.value, ld, 16, "fp64", stream=s.value)
    y_host = A.matvec(Q[:, 3], "fp32")  # null stream, after the device call
    assert hip.hipStreamSynchronize(s) == 0
    Yh = np.zeros_like(Bh)
    assert hip.hipMemcpy(Yh.ctypes.data, dY, Bh.nbytes, 2) == 0
    Ys = A.matvec(Q, "fp64")
    assert col_rel(Yh[:n], Ys) <= 1e-12
    assert rel(y_host, Ys[:, 3]) <= 1e-6
    hip.hipStreamDestroy(s)
    hip.hipFree(dB)
    hip.hipFree(dY)
    A.close()


@pytest.mark.parametrize("n,leaf,seed", [(30000, 64, 3), (8192, 32, 5)])
def test_fp64_near_pairs_ragged_leaves(engine, n, leaf, seed):
    """FP64 plans evaluate each neighbouring leaf pair once and add the
    transposed products through per-slot partials: leaves of 117 points (two
    row blocks, so two slots per lower neighbour) and of 32 (a partial row
    tile), against the single-RHS passes (which evaluate the near field in
    another kernel)."""
    print_case("near pairs", n, leaf, seed)
    xy, _ = uniform_points(n, seed)
    rng = np.random.default_rng(seed)
    Q = rng.uniform(-1, 1, size=(n, 6))
    A = engine.AFMM(n, leaf, 10, xy, kernel="log")
    A.assemble()
    cnt = A.NumberOfParticlesInLeaves
    print_case("  leaves %d, points per leaf %d..%d" % (len(cnt), cnt.min(), cnt.max()))
    Yb = A.matvec(Q, "fp64", blocked=True)
    Ys = A.matvec(Q, "fp64")
    print_case("  blocked vs single-RHS %.2e" % col_rel(Yb, Ys))
    assert col_rel(Yb, Ys) <= 1e-12
    A.close()


def test_fp64_near_pairs_match_both_sided_evaluation(engine, tmp_path):
    """The same blocked FP64 matvec with the near field's leaf pairs
    evaluated once (default) and from both sides (AFMM2D_MR_NEAR_PAIRS=0,
    read once per process: run in a child process)."""
    import os
    import subprocess
    import sys
    n = 20000
    script = tmp_path / "both_sides.py"
    out = tmp_path / "y.npy"
    script.write_text(
        "import sys, numpy as np\n"
        "sys.path.insert(0, %r)\n"
        "from conftest import uniform_points\n"
        "import afmm2d_amd\n"
        "xy, _ = uniform_points(%d, 21)\n"
        "Q = np.random.default_rng(22).uniform(-1, 1, size=(%d, 16))\n"
        "A = afmm2d_amd.AFMM(%d, 100, 10, xy, kernel='inv_r')\n"
        "A.assemble()\n"
        "np.save(%r, A.matvec(Q, 'fp64', blocked=True))\n"
        % (os.path.dirname(os.path.abspath(__file__)), n, n, n, str(out)))
    env = dict(os.environ, AFMM2D_MR_NEAR_PAIRS="0")
    r = subprocess.run([sys.executable, str(script)], env=env, capture_output=True, text=True,
                       timeout=120, cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert r.returncode == 0, r.stderr[-2000:]
    Y2 = np.load(out)
    xy, _ = uniform_points(n, 21)
    Q = np.random.default_rng(22).uniform(-1, 1, size=(n, 16))
    A = engine.AFMM(n, 100, 10, xy, kernel="inv_r")
    A.assemble()
    Y1 = A.matvec(Q, "fp64", blocked=True)
    print_case("near pairs vs both sides: %.2e" % col_rel(Y1, Y2))
    assert col_rel(Y1, Y2) <= 1e-13
    A.close()
