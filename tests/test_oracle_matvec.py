"""Pins of the oracle's operator (AFMM2D.cpp:3-74, AFMM2DTree.cpp:991-1236)
against the FP64 dense direct sum (tutorial.cpp:97-131 semantics) and the
identities any linear operator on a point set must satisfy."""
import numpy as np
import pytest

from conftest import uniform_points


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("kid,kp", [(0, None), (1, None), (3, None), (2, [0.05, 0, 0, 0]),
                                    (4, None), (4, [0.5, 0, 0, 0]), (4, [0.5, 0.25, 0, 0])])
def test_kernel_values(oracle_mod, kid, kp):
    rng = np.random.default_rng(0)
    a = 0.125
    for _ in range(50):
        xi, yi, xj, yj = rng.uniform(-1, 1, 4)
        r = np.hypot(xi - xj, yi - yj)
        # the example user kernel: anisotropic log plus a dipole term
        g = kp[0] if kp else 2.0
        b = kp[1] if kp else 0.0
        expect = {0: np.log(r), 1: 1 / r, 3: (r / a if r < a else a / r),
                  2: np.exp(-r * r / (kp[0] if kp else 0.01)),
                  4: np.log(np.hypot(xi - xj, g * (yi - yj))) + b * (xi - xj) / r ** 2}[kid]
        got = oracle_mod.kernel_value(kid, xi, yi, xj, yj, a=a, kparams=kp)
        assert abs(got - expect) <= 1e-14 * max(1.0, abs(expect))
    # diagonal values
    assert oracle_mod.kernel_value(kid, 0.3, 0.3, 0.3, 0.3, a=a, kparams=kp) == (1.0 if kid == 2 else 0.0)


def test_config1_error_vs_direct(oracle_mod):
    xy, q = uniform_points(4096, 1)
    T = oracle_mod.Tree(xy, 64, 10)
    assert T.L == 3
    T.assemble()
    y = T.matvec(q)
    e = rel(y, T.direct(q))
    assert e <= 1e-7, e
    # compute_error on every leaf agrees with the full comparison's scale
    errs = [T.compute_error(b) for b in range(64)]
    assert max(errs) <= 1e-6


def test_error_decreases_with_tolerance(oracle_mod):
    xy, q = uniform_points(16384, 2)
    errs = []
    for tp in (6, 9, 12):
        T = oracle_mod.Tree(xy, 64, tp)
        T.assemble()
        errs.append(rel(T.matvec(q), T.direct(q)))
        T.close()
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] <= 1e-8


def test_linearity_and_permutation_invariance(oracle_mod):
    xy, q1 = uniform_points(8192, 3)
    q2 = np.random.default_rng(4).uniform(-1, 1, 8192)
    T = oracle_mod.Tree(xy, 64, 8)
    T.assemble()
    y1, y2 = T.matvec(q1), T.matvec(q2)
    y3 = T.matvec(2.0 * q1 - 3.0 * q2)
    assert rel(y3, 2.0 * y1 - 3.0 * y2) <= 1e-13
    # relabelling the points relabels the output (the KD order depends only on
    # the coordinates), bit for bit
    p = np.random.default_rng(5).permutation(8192)
    T2 = oracle_mod.Tree(xy[p], 64, 8)
    T2.assemble()
    np.testing.assert_array_equal(T2.matvec(q1[p]), y1[p])


def test_rank_zero_far_field_is_exact(oracle_mod):
    # a very narrow Gaussian: every far-field entry underflows to 0, every ACA
    # returns rank 0 and the operator is the near field alone
    xy, q = uniform_points(4096, 6)
    T = oracle_mod.Tree(xy, 64, 10, kernel_id=2, kparams=[1e-5, 0, 0, 0])
    T.assemble()
    assert all((T.ranks(j) == 0).all() for j in range(2, T.L + 1))
    assert rel(T.matvec(q), T.direct(q)) <= 1e-14


def test_coincident_points(oracle_mod):
    # Duplicated points: K = 0 where R = 0, so every entry stays finite.  A
    # pair can straddle two KD boxes that the (combinatorial) interaction list
    # of the reference treats as far apart, so the far field near such a pair
    # is not low rank and the error can be large; the defined result is a
    # finite vector that matches between the oracle and the GPU path.
    xy, q = uniform_points(4096, 8)
    xy[100:200] = xy[0:100]
    T = oracle_mod.Tree(xy, 64, 10)
    T.assemble()
    y = T.matvec(q)
    assert np.isfinite(y).all()
    assert np.isfinite(T.direct(q)).all()


def test_parallel_modes_agree(oracle_mod):
    xy, q = uniform_points(16384, 9)
    T = oracle_mod.Tree(xy, 64, 8)
    T.assemble()
    T.set_parallel_mode(0)
    a = T.matvec(q)
    T.set_parallel_mode(1)
    b = T.matvec(q)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("kid,tol_pow", [(0, 12), (0, 8), (2, 10), (3, 6)])
def test_exact_skeleton_rows_change_the_matvec_by_rounding_only(oracle_mod, kid, tol_pow):
    # The product stores L2P without its skeleton rows and applies them as the
    # identity: row skel_q of Ac D^-1 is e_q in exact arithmetic.  The solved
    # rows differ from e_q by rounding amplified by cond(D); the operator must
    # not notice.
    xy, q = uniform_points(20000, 40 + kid)
    T = oracle_mod.Tree(xy, 64, tol_pow, kid)
    T.assemble()
    X = oracle_mod.Tree(xy, 64, tol_pow, kid)
    X.set_exact_skeleton_rows()
    X.assemble()
    for j in (2, T.L):
        for box in (0, 4 ** j // 2, 4 ** j - 1):
            P, Px = T.l2p(j, box), X.l2p(j, box)
            sk = T.skeleton(j, box)
            np.testing.assert_array_equal(sk, X.skeleton(j, box))
            if not P.size:
                continue
            # the rows of the box that are skeleton rows (as indices of the
            # box's own rows): leaf rows are its points, upper rows are the
            # children's skeletons in child order
            if j == T.L:
                st, cnt = T.leaf_range(box)
                rows = np.arange(st, st + cnt)
            else:
                rows = np.concatenate([T.skeleton(j + 1, 4 * box + c) for c in range(4)])
            pos = np.array([int(np.nonzero(rows == g)[0][0]) for g in sk])
            np.testing.assert_array_equal(Px[pos], np.eye(len(sk)))
            other = np.setdiff1d(np.arange(P.shape[0]), pos)
            np.testing.assert_array_equal(Px[other], P[other])
            assert np.abs(P[pos] - np.eye(len(sk))).max() < 1e-1
    y, yx, yd = T.matvec(q), X.matvec(q), T.direct(q)
    assert np.linalg.norm(yx - y) <= 1e-12 * np.linalg.norm(y)
    e, ex = (np.linalg.norm(v - yd) / np.linalg.norm(yd) for v in (y, yx))
    assert ex <= 1.01 * e + 1e-15
    T.close()
    X.close()


def test_invalid_levels(oracle_mod):
    xy, _ = uniform_points(100, 1)
    with pytest.raises(ValueError):
        oracle_mod.Tree(xy, 64, 8)


@pytest.mark.parametrize("kid", [0, 1, 2, 3, 4, 6])
def test_near_field_brute_force(oracle_mod, kid):
    """The oracle's near-field term (AFMM2DTree.cpp:1199-1213) against a
    numpy brute force over each leaf's own points and its 8 neighbours'."""
    xy, q = uniform_points(2000, 21)
    kp = [0.0, 0.25, 0.0, 0.0] if kid == 6 else None  # dipole part of the user kernel
    T = oracle_mod.Tree(xy, 30, 8, kid, kparams=kp)
    L = T.L
    assert L == 3
    sxy, perm = T.sorted_xy(), T.perm()
    qs = q[perm]
    nbr, _, _ = oracle_mod.topology(L)
    a = T.a
    ys = np.zeros(len(q))
    for k in range(4 ** L):
        st, n = T.leaf_range(k)
        boxes = [k] + [b for b in nbr[k] if b >= 0]
        for b in boxes:
            sb, m = T.leaf_range(b)
            d = sxy[st:st + n, None, :] - sxy[None, sb:sb + m, :]
            r2 = (d ** 2).sum(-1)
            with np.errstate(divide="ignore", invalid="ignore"):
                r = np.sqrt(r2)
                r2g = d[..., 0] ** 2 + (2.0 * d[..., 1]) ** 2
                K = {0: np.where(r2 > 0, 0.5 * np.log(r2), 0.0),
                     1: np.where(r2 > 0, 1.0 / r, 0.0),
                     2: np.exp(-r2 / 0.01),
                     3: np.where(r < a, r / a, a / r),
                     4: np.where(r2g > 0, 0.5 * np.log(r2g), 0.0),
                     6: np.where(r2 > 0, 0.25 * d[..., 0] / r2, 0.0)}[kid]
            ys[st:st + n] += K @ qs[sb:sb + m]
    y = np.empty_like(ys)
    y[perm] = ys
    got = T.near_field(q)
    assert rel(got, y) <= 1e-13


def test_user_kernel_operator(oracle_mod):
    """The compiled-in user kernel (id 4) through the whole oracle operator:
    error against the direct sum falls with the tolerance."""
    xy, q = uniform_points(4096, 31)
    errs = []
    for tp in (6, 10):
        T = oracle_mod.Tree(xy, 64, tp, 4)
        T.assemble()
        errs.append(rel(T.matvec(q), T.direct(q)))
        T.close()
    assert errs[1] < errs[0] and errs[1] <= 1e-7, errs


def test_user_kernel_parts(oracle_mod):
    """Kernel ids 5 and 6: the symmetric and antisymmetric parts of the user
    kernel, exactly (anti)symmetric and summing to it."""
    rng = np.random.default_rng(3)
    kp = [0.0, 0.25, 0.0, 0.0]
    for _ in range(100):
        xi, yi, xj, yj = rng.uniform(-1, 1, 4)
        k = oracle_mod.kernel_value(4, xi, yi, xj, yj, kparams=kp)
        kt = oracle_mod.kernel_value(4, xj, yj, xi, yi, kparams=kp)
        s_ij = oracle_mod.kernel_value(5, xi, yi, xj, yj, kparams=kp)
        a_ij = oracle_mod.kernel_value(6, xi, yi, xj, yj, kparams=kp)
        assert s_ij == oracle_mod.kernel_value(5, xj, yj, xi, yi, kparams=kp)
        assert a_ij == -oracle_mod.kernel_value(6, xj, yj, xi, yi, kparams=kp)
        assert abs(s_ij + a_ij - k) <= 1e-15 * max(1.0, abs(k))
        assert abs(a_ij - 0.25 * (xi - xj) / ((xi - xj) ** 2 + (yi - yj) ** 2)) <= \
            1e-12 * max(1.0, abs(a_ij))
        assert k != kt


def test_nonsymmetric_operator_as_two_parts(oracle_mod):
    """A non-symmetric kernel through the reference's algorithm, one operator
    per part: the sum approaches the direct sum of the full kernel as the
    tolerance falls (the antisymmetric part uses the same L2P^T multipole
    basis, since A(far, box) = -A(box, far)^T)."""
    kp = [0.0, 0.25, 0.0, 0.0]
    xy, q = uniform_points(4096, 37)
    T4 = oracle_mod.Tree(xy, 64, 6, 4, kparams=kp)
    yd = T4.direct(q)
    errs = []
    for tp in (6, 10):
        y = np.zeros_like(q)
        for kid in (5, 6):
            T = oracle_mod.Tree(xy, 64, tp, kid, kparams=kp)
            T.assemble()
            y += T.matvec(q)
            T.close()
        errs.append(rel(y, yd))
    # the near field of the full kernel is the two parts' near fields
    ns = oracle_mod.Tree(xy, 64, 6, 5, kparams=kp).near_field(q)
    na = oracle_mod.Tree(xy, 64, 6, 6, kparams=kp).near_field(q)
    assert rel(ns + na, T4.near_field(q)) <= 1e-14
    assert errs[1] < errs[0] and errs[1] <= 1e-7, errs
