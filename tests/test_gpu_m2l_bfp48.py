"""48-bit block floating-point M2L storage (afmm2d_set_m2l_storage(h,
AFMM2D_M2L_BFP48)): every dense interaction block kept as 48-bit integer
mantissas with one exponent per block, the products still formed in FP64.

Pinned two ways: the stored blocks equal the FP64 assembly's blocks rounded
by bfp48_round (itself pinned on the CPU by what the format means), bit for
bit, on every level and in both orientations; the matvec stays
within 1e-11 of the FP64-block matvec and of the oracle (the FP64 path's
own bar), for every built-in kernel, the streamed (rank > 64) classes, the
user kernel declared symmetric and a non-symmetric one (two operators), and
with ACA-factored pairs beside the rounded dense ones."""
import numpy as np
import pytest

from conftest import bfp48_round, oracle_kernel, uniform_points

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def pair(engine, n, leaf, tol_pow, xy, kernel, kp=None, factoring=False):
    out = []
    for storage in ("fp64", "bfp48"):
        a = engine.AFMM(n, leaf, tol_pow, xy, kernel=kernel, kparams=kp)
        a.set_m2l_storage(storage)
        if factoring:
            a.set_m2l_factoring(True)
        a.assemble()
        out.append(a)
    return out


@pytest.mark.parametrize("n,leaf,tol_pow,kernel", [
    (20000, 64, 12, "log"), (10000, 100, 12, "reg_inv_r")])
def test_stored_blocks_are_the_rounded_fp64_blocks(engine, n, leaf, tol_pow, kernel):
    print("bfp48 blocks", n, leaf, tol_pow, kernel, flush=True)
    xy, _ = uniform_points(n, 3 + n)
    A, C = pair(engine, n, leaf, tol_pow, xy, kernel)
    sa, sc = A.stats(), C.stats()
    assert sa["m2l_storage"] == 0 and sc["m2l_storage"] == 1
    assert 4 * sc["m2l_bytes"] == 3 * sa["m2l_bytes"]
    rng = np.random.default_rng(n)
    nblk = nchanged = 0
    for j in range(2, A.nLevels + 1):
        np.testing.assert_array_equal(A.ranks(j), C.ranks(j))
        for box in rng.choice(4 ** j, size=min(4 ** j, 16), replace=False):
            for slot in range(40):
                a = A.m2l(j, box, slot)
                if a is None or a.size == 0:
                    continue
                c = C.m2l(j, box, slot)
                np.testing.assert_array_equal(c, bfp48_round(a)[0])
                nblk += 1
                nchanged += int(np.any(c != a))
    print("  %d sampled blocks bit-equal to the rounded FP64 blocks (%d changed by the rounding)"
          % (nblk, nchanged), flush=True)
    assert nblk > 0 and nchanged > 0
    A.close()
    C.close()


@pytest.mark.parametrize("n,leaf,tol_pow,kernel", [
    (20000, 64, 12, "log"), (8192, 64, 10, "inv_r"), (8192, 64, 6, "gaussian"),
    (8192, 64, 10, "reg_inv_r"),
    # ranks above 64: the streamed M2L classes
    (10000, 100, 12, "reg_inv_r"), (5000, 100, 14, "reg_inv_r"),
    (20000, 64, 12, "user_symmetric")])
def test_matvec_matches_fp64_blocks_and_oracle(engine, oracle_mod, n, leaf, tol_pow, kernel):
    print("bfp48 matvec", n, leaf, tol_pow, kernel, flush=True)
    xy, q = uniform_points(n, 11 + tol_pow)
    A, C = pair(engine, n, leaf, tol_pow, xy, kernel)
    if leaf == 100:
        assert C.stats()["max_rank"] > 64
    T = oracle_mod.Tree(xy, leaf, tol_pow, oracle_kernel(kernel))
    T.assemble()
    ya, yc, yo = A.computeMatVecProduct(q), C.computeMatVecProduct(q), T.matvec(q)
    d_a, d_o = rel(yc, ya), rel(yc, yo)
    print("  rel diff vs FP64 blocks %.2e, vs oracle %.2e (FP64 blocks vs oracle %.2e)"
          % (d_a, d_o, rel(ya, yo)), flush=True)
    assert d_a <= 1e-11
    assert d_o <= 1e-11
    T.close()
    A.close()
    C.close()


def test_nonsymmetric_user_kernel(engine):
    kp = [0.0, 0.25, 0.0, 0.0]
    n = 12000
    xy, q = uniform_points(n, 95)
    A, C = pair(engine, n, 64, 10, xy, "user", kp)
    ya, yc = A.computeMatVecProduct(q), C.computeMatVecProduct(q)
    print("bfp48 non-symmetric user kernel: rel diff %.2e" % rel(yc, ya), flush=True)
    assert rel(yc, ya) <= 1e-11
    rows = np.arange(0, n, 53)
    err_a, err_c = rel(ya[rows], A.direct_rows(rows, q)), rel(yc[rows], A.direct_rows(rows, q))
    assert abs(err_c - err_a) <= 1e-11
    A.close()
    C.close()


def test_with_factored_pairs(engine):
    n = 20000
    xy, q = uniform_points(n, 97)
    A, C = pair(engine, n, 64, 10, xy, "log", factoring=True)
    sa, sc = A.stats(), C.stats()
    assert sc["m2l_factored_pairs"] == sa["m2l_factored_pairs"] > 0
    assert sc["m2l_factored_bytes"] == sa["m2l_factored_bytes"]
    dense_a = sa["m2l_bytes"] - sa["m2l_factored_bytes"]
    assert sc["m2l_bytes"] == 3 * dense_a // 4 + sc["m2l_factored_bytes"]
    ya, yc = A.computeMatVecProduct(q), C.computeMatVecProduct(q)
    print("bfp48 with factored pairs: rel diff %.2e" % rel(yc, ya), flush=True)
    assert rel(yc, ya) <= 1e-11
    # factored pairs keep their FP64 factors; dense ones are rounded
    L = C.nLevels
    nf = nd = 0
    for box in range(0, 4 ** L, 97):
        for slot in range(40):
            where = C.m2l_storage(L, box, slot)
            if where is None:
                continue
            a, c = A.m2l(L, box, slot), C.m2l(L, box, slot)
            if where[0] >= 0:
                np.testing.assert_array_equal(c, a)
                nf += 1
            else:
                np.testing.assert_array_equal(c, bfp48_round(a)[0])
                nd += 1
    assert nf > 0 and nd > 0
    A.close()
    C.close()


def test_storage_is_set_before_assembly(engine):
    xy, _ = uniform_points(4096, 1)
    A = engine.AFMM(4096, 64, 10, xy)
    with pytest.raises(ValueError):
        A.set_m2l_storage("fp32")
    A.assemble()
    with pytest.raises(engine.AFMM2DError) as ex:
        A.set_m2l_storage("bfp48")
    assert ex.value.status == engine.ERR_STATE
    A.close()


@pytest.mark.parametrize("n,leaf,tol_pow,kernel,nrhs", [
    (20000, 64, 12, "log", 16), (20000, 64, 12, "log", 32), (8192, 64, 10, "inv_r", 5),
    # ranks above 64 (long interaction-list stacks: descriptors past the LDS budget)
    (10000, 100, 12, "reg_inv_r", 16), (5000, 100, 14, "reg_inv_r", 3),
    (8192, 64, 10, "gaussian", 8)])
def test_blocked_passes_use_the_same_blocks(engine, oracle_mod, n, leaf, tol_pow, kernel, nrhs):
    """On a handle with 48-bit M2L storage the blocked FP64 passes read every
    pair's block and its transpose in the same 48-bit form: the same stored
    values as the single-RHS passes, so the two agree like the FP64 forms do
    (1e-12), and both stay within 1e-11 of the oracle."""
    print("bfp48 blocked", n, leaf, tol_pow, kernel, nrhs, flush=True)
    xy, _ = uniform_points(n, 31 + n)
    kp = [0.01, 0, 0, 0] if kernel == "gaussian" else None
    A = engine.AFMM(n, leaf, tol_pow, xy, kernel=kernel, kparams=kp)
    A.set_m2l_storage("bfp48")
    A.assemble()
    Q = np.random.default_rng(n).uniform(-1, 1, size=(n, nrhs))
    Yb = A.matvec(Q, "fp64", blocked=True)
    Ys = A.matvec(Q, "fp64")
    d = max(rel(Yb[:, c], Ys[:, c]) for c in range(nrhs))
    T = oracle_mod.Tree(xy, leaf, tol_pow, oracle_kernel(kernel), kparams=kp)
    T.assemble()
    do = max(rel(Yb[:, c], T.matvec(Q[:, c])) for c in range(min(nrhs, 2)))
    print("  blocked vs single-RHS %.2e, vs oracle %.2e" % (d, do), flush=True)
    assert d <= 1e-12
    assert do <= 1e-11
    T.close()
    A.close()


def test_blocked_nonsymmetric_user_kernel(engine):
    kp = [0.0, 0.25, 0.0, 0.0]
    n = 12000
    xy, _ = uniform_points(n, 96)
    A = engine.AFMM(n, 64, 10, xy, kernel="user", kparams=kp)
    A.set_m2l_storage("bfp48")
    A.assemble()
    Q = np.random.default_rng(97).uniform(-1, 1, size=(n, 6))
    Yb = A.matvec(Q, "fp64", blocked=True)
    Ys = A.matvec(Q, "fp64")
    d = max(rel(Yb[:, c], Ys[:, c]) for c in range(6))
    print("bfp48 blocked, non-symmetric user kernel: vs single-RHS %.2e" % d, flush=True)
    assert d <= 1e-12
    A.close()
