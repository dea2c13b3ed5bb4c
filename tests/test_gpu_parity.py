"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs."""
import numpy as np
import pytest

from conftest import uniform_points, oracle_kernel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,leaf,seed,quant", [
    (4096, 64, 1, 0), (5000, 64, 2, 0), (70001, 100, 3, 0), (20000, 64, 4, 64), (3000, 40, 5, 8)])
def test_kd_partition_matches_oracle(engine, oracle_mod, n, leaf, seed, quant):
    xy, _ = uniform_points(n, seed)
    if quant:  # many equal coordinates: exercises the reference's tie order
        xy = np.round(xy * quant) / quant
    L = oracle_mod.nlevels(n, leaf)
    sxy, perm, counts, boxnum = oracle_mod.kd_sort(xy, L)
    a = engine.AFMM(n, leaf, 8, xy)
    assert a.nLevels == L
    np.testing.assert_array_equal(a.sorted_Properties, perm)
    np.testing.assert_array_equal(a.sorted_Locations, sxy)
    np.testing.assert_array_equal(a.NumberOfParticlesInLeaves, counts)
    a.close()


def test_topology_matches_oracle(engine, oracle_mod):
    xy, _ = uniform_points(4 ** 5 * 4, 7)
    a = engine.AFMM(xy.shape[0], 4, 8, xy)
    assert a.nLevels == 5
    for level in range(0, 6):
        got = a.interactions(level)
        ref = oracle_mod.topology(level)
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)
    a.close()


@pytest.mark.parametrize("n,leaf,tol_pow,kernel", [
    (4096, 64, 10, "log"), (4096, 64, 8, "log"), (16384, 64, 12, "log"),
    (8192, 64, 10, "inv_r"), (8192, 64, 10, "reg_inv_r"), (8192, 64, 6, "gaussian"),
    # ranks above 64: the streamed M2L classes (leaf level and upper levels)
    (10000, 100, 12, "reg_inv_r"), (5000, 100, 14, "reg_inv_r"),
    # the compiled-in user kernel (anisotropic log), declared symmetric
    (8192, 64, 10, "user_symmetric"), (20000, 64, 12, "user_symmetric")])
def test_assembly_and_matvec_match_oracle(engine, oracle_mod, n, leaf, tol_pow, kernel):
    kid = oracle_kernel(kernel)
    xy, q = uniform_points(n, 11 + tol_pow)
    T = oracle_mod.Tree(xy, leaf, tol_pow, kid)
    T.assemble()
    a = engine.AFMM(n, leaf, tol_pow, xy, kernel=kernel)
    a.assemble()
    if leaf == 100:
        assert a.stats()["max_rank"] > 64  # blocks beyond the register tiles
    mism = 0
    tot = 0
    for j in range(2, T.L + 1):
        rg, ro = a.ranks(j), T.ranks(j)
        mism += int((rg != ro).sum())
        tot += rg.size
    assert mism == 0, "rank mismatches %d of %d" % (mism, tot)
    for j in (2, T.L):
        for box in (0, 4 ** j // 3, 4 ** j - 1):
            sk = a.skeleton(j, box)
            np.testing.assert_array_equal(sk, T.skeleton(j, box))
            P, Po = a.l2p(j, box), T.l2p(j, box)
            assert P.shape == Po.shape
            if P.size:
                # L2P = Ac D^{-1}: 1-ulp differences of the kernel values are
                # amplified by cond(D), so the coefficients agree to a
                # relative 1e-4 while the operator they define agrees far
                # tighter (the matvec check below).
                assert np.linalg.norm(P - Po) <= 1e-4 * np.linalg.norm(Po)
                # skeleton rows are stored as exactly the identity
                if j == T.L:
                    st, cnt = T.leaf_range(box)
                    rows = np.arange(st, st + cnt)
                else:
                    rows = np.concatenate([T.skeleton(j + 1, 4 * box + c) for c in range(4)])
                pos = [int(np.nonzero(rows == g)[0][0]) for g in sk]
                np.testing.assert_array_equal(P[pos], np.eye(len(sk)))
    y = a.computeMatVecProduct(q)
    yo = T.matvec(q)
    rel = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    print("matvec rel diff vs oracle %.3e" % rel)
    assert rel <= 1e-11, rel
    T.close()
    a.close()


def test_direct_sum_matches_oracle(engine, oracle_mod):
    xy, q = uniform_points(4096, 21)
    T = oracle_mod.Tree(xy, 64, 10)
    a = engine.AFMM(4096, 64, 10, xy)
    yd = a.direct(q)
    yo = T.direct(q)
    assert np.linalg.norm(yd - yo) / np.linalg.norm(yo) <= 1e-13
    T.close()
    a.close()


@pytest.mark.parametrize("n,leaf,kernel,quant", [
    (3000, 40, "log", 0),          # 47 points per leaf: LDS arrays of 64
    (7000, 100, "reg_inv_r", 0),   # 109: 128
    (20000, 64, "gaussian", 0),    # 78: 128
    (12000, 64, "log", 0),         # 187: 256
    (12000, 64, "reg_inv_r", 0),
    (5000, 100, "inv_r", 0),       # 312: 512
    (5000, 100, "reg_inv_r", 16),  # duplicated points: R = 0 off the diagonal
    (6000, 64, "inv_r", 16),
    (7000, 100, "user_symmetric", 0),  # user kernel: padding repeats a real point
    (6000, 64, "user_symmetric", 16),
])
def test_near_field_matches_oracle(engine, oracle_mod, n, leaf, kernel, quant):
    """The near-field kernel alone (AFMM2DTree.cpp:1199-1213) against the
    oracle's: the rsqrt-based 1/R kernels, the table log and exp agree to a few
    ulp per entry, so the sums agree to 1e-13 relative."""
    kid = oracle_kernel(kernel)
    xy, q = uniform_points(n, 40 + n % 97)
    if quant:
        xy = np.round(xy * quant) / quant
    T = oracle_mod.Tree(xy, leaf, 8, kid)
    ref = T.near_field(q)
    a = engine.AFMM(n, leaf, 8, xy, kernel=kernel)
    got = a.near_field(q)
    a.close()
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    assert err <= 1e-13, err
    assert float(np.max(np.abs(got - ref))) <= 1e-12 * float(np.max(np.abs(ref)))


@pytest.mark.parametrize("scale,quant", [(1e-7, 0), (1e3, 0), (1.0, 16), (1.0, 4096)])
def test_log_near_field_over_scales(engine, oracle_mod, scale, quant):
    """The log kernel's table evaluation (log2 R^2 from a 256-entry table and
    a degree-5 polynomial, ln(2)/2 applied to the sums) against the oracle's
    0.5 log(R^2) over squared distances from ~1e-16 to ~1e7; duplicated
    points (quant) give R = 0 off the diagonal, which must add exactly
    nothing."""
    n = 6000
    xy, q = uniform_points(n, 91)
    if quant:
        xy = np.round(xy * quant) / quant
    xy = xy * scale
    T = oracle_mod.Tree(xy, 64, 8, oracle_kernel("log"))
    ref = T.near_field(q)
    a = engine.AFMM(n, 64, 8, xy, kernel="log")
    got = a.near_field(q)
    a.close()
    T.close()
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print("log near field, scale %g, quant %d: rel %.2e" % (scale, quant, err), flush=True)
    assert err <= 1e-13, err
    assert float(np.max(np.abs(got - ref))) <= 1e-12 * float(np.max(np.abs(ref)))


@pytest.mark.parametrize("p0", [1e-6, 1e-2, 1e2])
def test_gaussian_near_field_over_widths(engine, oracle_mod, p0):
    """The Gaussian's table evaluation (exp by a 64-entry 2^(j/64) table and
    a degree-5 polynomial) against the oracle's exp(-R^2/p0), from widths
    where most values underflow to widths where all are near 1."""
    n = 20000
    xy, q = uniform_points(n, 93)
    kp = [p0, 0.0, 0.0, 0.0]
    T = oracle_mod.Tree(xy, 64, 8, oracle_kernel("gaussian"), kparams=kp)
    ref = T.near_field(q)
    a = engine.AFMM(n, 64, 8, xy, kernel="gaussian", kparams=kp)
    got = a.near_field(q)
    a.close()
    T.close()
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print("gaussian near field, p0 %g: rel %.2e" % (p0, err), flush=True)
    assert err <= 1e-13, err
    assert float(np.max(np.abs(got - ref))) <= 1e-12 * float(np.max(np.abs(ref)))


@pytest.mark.parametrize("n,leaf,kernel,chunks", [
    (14336, 200, "log", 4),        # 224 points per leaf: four 64-point chunks
    (14336, 200, "reg_inv_r", 4),
    (7680, 400, "inv_r", 8),       # 480: eight chunks, two groups of four
    (14336, 200, "user_symmetric", 4),
    (20000, 300, "log", 5),        # 312: chunks two at a time, then one
    (100000, 100, "reg_inv_r", 7),  # 390
])
def test_near_field_long_leaves_one_wave_per_pair(engine, oracle_mod, monkeypatch, n, leaf,
                                                  kernel, chunks):
    # One workgroup per leaf pair (P2P_WAVES = 0), so a wave holds every A
    # chunk of a leaf: four at a time when every chunk is full, else two at
    # a time and one; near field and matvec against the oracle.
    monkeypatch.setenv("AFMM2D_P2P_WAVES", "0")
    kid = oracle_kernel(kernel)
    xy, q = uniform_points(n, 5 + n % 89)
    T = oracle_mod.Tree(xy, leaf, 10, kid)
    a = engine.AFMM(n, leaf, 10, xy, kernel=kernel)
    counts = np.asarray(a.NumberOfParticlesInLeaves)
    assert ((counts + 63) // 64).min() == ((counts + 63) // 64).max() == chunks
    ref, got = T.near_field(q), a.near_field(q)
    assert float(np.linalg.norm(got - ref) / np.linalg.norm(ref)) <= 1e-13
    T.assemble()
    a.assemble()
    y, yo = a.computeMatVecProduct(q), T.matvec(q)
    assert float(np.linalg.norm(y - yo) / np.linalg.norm(yo)) <= 1e-11
    a.close()
    T.close()


def test_user_kernel_parameters_reach_the_device(engine, oracle_mod):
    """kparams of the user kernel (here the anisotropy g = 3) used by the
    device kernels: near field and whole operator against the oracle."""
    xy, q = uniform_points(8192, 77)
    kp = [3.0, 0.0, 0.0, 0.0]
    T = oracle_mod.Tree(xy, 64, 10, 4, kparams=kp)
    a = engine.AFMM(8192, 64, 10, xy, kernel="user_symmetric", kparams=kp)
    nf, nf_ref = a.near_field(q), T.near_field(q)
    assert np.linalg.norm(nf - nf_ref) <= 1e-13 * np.linalg.norm(nf_ref)
    T.assemble()
    a.assemble()
    y, yo = a.computeMatVecProduct(q), T.matvec(q)
    assert np.linalg.norm(y - yo) <= 1e-11 * np.linalg.norm(yo)
    # a different g is a different operator
    a2 = engine.AFMM(8192, 64, 10, xy, kernel="user_symmetric")
    assert np.linalg.norm(a2.near_field(q) - nf) > 1e-3 * np.linalg.norm(nf)
    a.close()
    a2.close()
    T.close()


@pytest.mark.parametrize("tol_pow", [8, 11])
def test_nonsymmetric_user_kernel(engine, oracle_mod, tol_pow):
    """A non-symmetric user kernel (the example with a dipole term) runs as
    its symmetric and antisymmetric parts: each part against the oracle
    (exact ranks, matvec to 1e-11), the whole operator against their sum,
    the direct sum of the full kernel and the oracle's full near field."""
    kp = [0.0, 0.25, 0.0, 0.0]
    n = 12000
    xy, q = uniform_points(n, 91 + tol_pow)
    parts = np.zeros(n)
    for name, kid in (("user_sym", 5), ("user_anti", 6)):
        T = oracle_mod.Tree(xy, 64, tol_pow, kid, kparams=kp)
        T.assemble()
        a = engine.AFMM(n, 64, tol_pow, xy, kernel=name, kparams=kp)
        a.assemble()
        for j in range(2, T.L + 1):
            np.testing.assert_array_equal(a.ranks(j), T.ranks(j))
        y, yo = a.computeMatVecProduct(q), T.matvec(q)
        assert np.linalg.norm(y - yo) <= 1e-11 * np.linalg.norm(yo), name
        parts += yo
        a.close()
        T.close()
    a = engine.AFMM(n, 64, tol_pow, xy, kernel="user", kparams=kp)
    a.assemble()
    y = a.computeMatVecProduct(q)
    assert np.linalg.norm(y - parts) <= 1e-11 * np.linalg.norm(parts)
    T4 = oracle_mod.Tree(xy, 64, tol_pow, 4, kparams=kp)
    yd = a.direct(q)
    rows = np.arange(0, n, 97)
    np.testing.assert_allclose(yd[rows], T4.direct(q)[rows], rtol=0, atol=1e-13 * np.abs(yd).max())
    err = np.linalg.norm(y - yd) / np.linalg.norm(yd)
    assert err <= (1e-4 if tol_pow == 8 else 1e-8), err
    nf, nf_ref = a.near_field(q), T4.near_field(q)
    assert np.linalg.norm(nf - nf_ref) <= 1e-13 * np.linalg.norm(nf_ref)
    assert a.getMatrixEntry(0, 1) != a.getMatrixEntry(1, 0)
    a.close()
    T4.close()


def test_user_kernel_symmetry_is_declared(engine, oracle_mod):
    """An undeclared user kernel runs as its symmetric plus antisymmetric
    parts, whatever a sample suggests; the declared-symmetric id runs it as
    one operator.  For a symmetric kernel both give the oracle's operator;
    afmm2d_check_symmetry tells the two kernels apart."""
    n = 12000
    xy, q = uniform_points(n, 123)
    T = oracle_mod.Tree(xy, 64, 10, 4)
    T.assemble()
    yo = T.matvec(q)
    split = engine.AFMM(n, 64, 10, xy, kernel="user")
    decl = engine.AFMM(n, 64, 10, xy, kernel="user_symmetric")
    split.assemble()
    decl.assemble()
    assert split.check_symmetry() and decl.check_symmetry()
    for a in (split, decl):
        y = a.computeMatVecProduct(q)
        assert np.linalg.norm(y - yo) <= 1e-11 * np.linalg.norm(yo)
    asym = engine.AFMM(n, 64, 10, xy, kernel="user_symmetric", kparams=[0.0, 0.25, 0, 0])
    assert not asym.check_symmetry()
    for a in (split, decl, asym):
        a.close()
    T.close()
