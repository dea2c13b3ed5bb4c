"""Pins of the oracle's KD partition (KDTree.cpp:37-153, 200-243).

The oracle restates the reference's merge sort (ties take the right run).  It
is checked here against an independent numpy statement of the same order --
sort by (key ascending, position descending) -- on inputs with many ties, and
against the split invariants every KD box must satisfy.
"""
import numpy as np
import pytest

from conftest import uniform_points


def kd_brute(xy, L):
    counts = []

    def sort_seg(seg, key):
        pos = np.arange(len(seg))
        return seg[np.lexsort((-pos, xy[seg, key]))]

    def rec(seg, lev):
        seg = sort_seg(seg, 0)
        nl = len(seg) // 2
        left, right = sort_seg(seg[:nl], 1), sort_seg(seg[nl:], 1)
        lb, rb = len(left) // 2, len(right) // 2
        parts = [left[:lb], left[lb:], right[:rb], right[rb:]]
        if lev < L:
            lev += 1
            if lev == L:
                counts.extend(len(p) for p in parts)
            parts = [rec(p, lev) for p in parts]
        return np.concatenate(parts)

    perm = rec(np.arange(len(xy)), 0)
    return perm, np.array(counts)


@pytest.mark.parametrize("n,L,seed,quant", [
    (256, 2, 1, 0), (1000, 2, 2, 0), (4096, 3, 3, 0), (4096, 3, 4, 4), (3001, 3, 5, 16),
    (5000, 4, 6, 2), (777, 2, 7, 1)])
def test_kd_matches_bruteforce(oracle_mod, n, L, seed, quant):
    xy, _ = uniform_points(n, seed)
    if quant:
        xy = np.round(xy * quant) / quant
    sxy, perm, counts, boxnum = oracle_mod.kd_sort(xy, L)
    bperm, bcounts = kd_brute(xy, L)
    np.testing.assert_array_equal(perm, bperm)
    np.testing.assert_array_equal(counts, bcounts)
    np.testing.assert_array_equal(sxy, xy[perm])


def test_merge_sort_tie_order_is_reversed(oracle_mod):
    # All points equal: every sort reverses its input, so the final order is
    # a pure function of the split sizes; the brute force predicts it.
    xy = np.zeros((64, 2))
    _, perm, _, _ = oracle_mod.kd_sort(xy, 2)
    np.testing.assert_array_equal(perm, kd_brute(xy, 2)[0])
    assert sorted(perm.tolist()) == list(range(64))


def test_kd_invariants(oracle_mod):
    n, L = 10000, 4
    xy, _ = uniform_points(n, 9)
    sxy, perm, counts, boxnum = oracle_mod.kd_sort(xy, L)
    assert sorted(perm.tolist()) == list(range(n))
    assert counts.sum() == n
    # sizes follow floor/ceil halving
    sizes = [n]
    for _ in range(L):
        nxt = []
        for m in sizes:
            nl = m // 2
            nr = m - nl
            nxt += [nl // 2, nl - nl // 2, nr // 2, nr - nr // 2]
        sizes = nxt
    np.testing.assert_array_equal(counts, sizes)
    # every level-d box splits by x then each half by y
    off = np.concatenate([[0], np.cumsum(counts)])
    for d in range(L):
        per = 4 ** (L - d)
        for b in range(4 ** d):
            lo, hi = off[b * per], off[(b + 1) * per]
            q = per // 4
            c = [off[b * per + i * q] for i in range(5)]
            LB, LT, RB, RT = (sxy[c[i]:c[i + 1]] for i in range(4))
            left, right = sxy[c[0]:c[2]], sxy[c[2]:c[4]]
            assert left[:, 0].max() <= right[:, 0].min()
            assert LB[:, 1].max() <= LT[:, 1].min()
            assert RB[:, 1].max() <= RT[:, 1].min()
            assert hi - lo == c[4] - c[0]
    # KD-order leaf t is box number digits(t) mapped through [0,3,1,2]
    m = [0, 3, 1, 2]
    for t in (0, 1, 2, 3, 4, 77, 255):
        b = 0
        for d in range(L - 1, -1, -1):
            b = 4 * b + m[(t >> (2 * d)) & 3]
        assert boxnum[t] == b


@pytest.mark.parametrize("n,m,L", [(4096, 64, 3), (1 << 20, 64, 7), (1 << 22, 64, 8),
                                   (1 << 23, 128, 8), (6400, 200, 2), (1000, 100, 1),
                                   (100, 100, 0), (65536, 64, 5), (1000000, 100, 6)])
def test_nlevels(oracle_mod, n, m, L):
    assert oracle_mod.nlevels(n, m) == L
