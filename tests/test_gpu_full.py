"""Full-size checks at N = 2^20 (leaf 64, L = 7, log kernel): the benchmark
workload.  GPU vs the CPU oracle on the same operator, and the error bar of
the metric (relative 2-norm <= 1e-8 against the FP64 direct sum) at
TOL_POW = 12, checked on every row."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full():
    rng = np.random.default_rng(2024)
    n = 1 << 20
    xy = rng.uniform(-1.0, 1.0, size=(n, 2))
    q = rng.uniform(-1.0, 1.0, size=n)
    return n, xy, q


def test_full_size_matches_oracle_and_direct_sum(engine, oracle_mod, full):
    n, xy, q = full
    A = engine.AFMM(n, 64, 12, xy)
    A.assemble()
    assert A.nLevels == 7
    y = A.computeMatVecProduct(q)
    yd = A.direct(q)                      # GPU FP64 direct sum, all N rows
    err = float(np.linalg.norm(y - yd) / np.linalg.norm(yd))
    print("full N=2^20 tol 1e-12: rel.err vs direct %.3e" % err)
    assert err <= 1e-8
    # the direct sum itself against the CPU oracle's on sampled rows
    rows = np.sort(np.random.default_rng(1).choice(n, 256, replace=False))
    T = oracle_mod.Tree(xy, 64, 12)
    perm = T.perm()
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    d_cpu = T.direct_rows(inv[rows].astype(np.int32), q[perm])
    assert float(np.linalg.norm(yd[rows] - d_cpu) / np.linalg.norm(d_cpu)) <= 1e-13
    # and the whole operator against the oracle's.  The ACA's stopping test
    # compares norm estimates built from dot products; the oracle sums them
    # in ascending order, the engine in a wave's order, so a borderline
    # |prev - normS| >= tol prev can go either way on a few boxes
    T.assemble()
    mism = [int((A.ranks(j) != T.ranks(j)).sum()) for j in range(2, 8)]
    print("full N=2^20 tol 1e-12: rank mismatches per level 2..7 (ascending sums): %s" % mism)
    assert sum(mism) <= 4, mism   # observed: [0, 0, 0, 0, 0, 2]
    yo = T.matvec(q)
    d = float(np.linalg.norm(y - yo) / np.linalg.norm(yo))
    print("full N=2^20: GPU vs oracle rel.diff %.3e" % d)
    assert d <= 1e-11
    T.close()
    # with the engine's summation order the oracle's ranks move towards the
    # engine's; what is left comes from the kernel values themselves (the
    # device log differs from the host's in the last bit on some entries)
    T2 = wave_order_oracle(oracle_mod, xy, 12)
    mism2 = [int((A.ranks(j) != T2.ranks(j)).sum()) for j in range(2, 8)]
    d2 = float(np.linalg.norm(y - T2.matvec(q)) / np.linalg.norm(yo))
    print("full N=2^20 tol 1e-12: rank mismatches per level 2..7 (wave-order sums): %s, "
          "rel.diff %.3e" % (mism2, d2))
    assert sum(mism2) <= sum(mism)
    assert d2 <= 1e-11
    T2.close()
    A.close()


def wave_order_oracle(oracle_mod, xy, tol_pow):
    oracle_mod.set_aca_wave_sums(True)
    try:
        T = oracle_mod.Tree(xy, 64, tol_pow)
        T.assemble()
    finally:
        oracle_mod.set_aca_wave_sums(False)
    return T


def test_full_size_configured_tolerance(engine, oracle_mod, full):
    # the configured ACA tolerance 1e-8 gives ~4e-6 (the reference's method
    # at that tolerance): reported and bounded against the direct sum; the
    # operator itself against the oracle's at the same tolerance
    n, xy, q = full
    A = engine.AFMM(n, 64, 8, xy)
    A.assemble()
    y = A.computeMatVecProduct(q)
    rows = np.sort(np.random.default_rng(2).choice(n, 4096, replace=False))
    ref = A.direct_rows(rows, q)
    err = float(np.linalg.norm(y[rows] - ref) / np.linalg.norm(ref))
    print("full N=2^20 tol 1e-8: rel.err vs direct (4096 rows) %.3e" % err)
    assert err <= 1e-4
    T = oracle_mod.Tree(xy, 64, 8)
    T.assemble()
    mism = [int((A.ranks(j) != T.ranks(j)).sum()) for j in range(2, 8)]
    print("full N=2^20 tol 1e-8: rank mismatches per level 2..7 (ascending sums): %s" % mism)
    assert sum(mism) <= 8, mism
    yo = T.matvec(q)
    d = float(np.linalg.norm(y - yo) / np.linalg.norm(yo))
    print("full N=2^20 tol 1e-8: GPU vs oracle rel.diff %.3e" % d)
    # a differing rank moves the result by about the ACA error
    assert d <= (1e-11 if sum(mism) == 0 else 1e-6)
    T.close()
    T2 = wave_order_oracle(oracle_mod, xy, 8)
    mism2 = [int((A.ranks(j) != T2.ranks(j)).sum()) for j in range(2, 8)]
    d2 = float(np.linalg.norm(y - T2.matvec(q)) / np.linalg.norm(yo))
    print("full N=2^20 tol 1e-8: rank mismatches per level 2..7 (wave-order sums): %s, "
          "rel.diff %.3e" % (mism2, d2))
    assert sum(mism2) <= sum(mism)
    assert d2 <= (1e-11 if sum(mism2) == 0 else 1e-6)
    T2.close()
    A.close()


def test_published_shape_error_bar(engine):
    # The reference's published N = 1e6 shape (leaf 100: 244-point leaves,
    # ranks to 105, streamed M2L blocks, the near field four chunks at a
    # time), a/R kernel, tolerance 1e-12: error against the FP64 direct sum
    # on sampled rows.
    n = 1000000
    rng = np.random.default_rng(2024)
    xy = rng.uniform(-1.0, 1.0, size=(n, 2))
    q = rng.uniform(-1.0, 1.0, size=n)
    A = engine.AFMM(n, 100, 12, xy, kernel="reg_inv_r")
    A.assemble()
    assert A.nLevels == 6 and A.stats()["max_rank"] > 96
    y = A.computeMatVecProduct(q)
    rows = np.sort(np.random.default_rng(3).choice(n, 4096, replace=False))
    ref = A.direct_rows(rows, q)
    err = float(np.linalg.norm(y[rows] - ref) / np.linalg.norm(ref))
    print("N=1e6 leaf 100 a/R tol 1e-12: rel.err vs direct (4096 rows) %.3e" % err)
    assert err <= 1e-7
    A.close()
