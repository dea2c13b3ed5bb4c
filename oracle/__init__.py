"""ctypes binding of the CPU oracle (afmm2d_oracle.c).

TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module.  The product package (afmm2d_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libafmm2d_oracle.so")
_lib = None

KERNEL_LOG, KERNEL_INV_R, KERNEL_GAUSS, KERNEL_REG_INV_R, KERNEL_USER = 0, 1, 2, 3, 4
KERNEL_USER_SYM, KERNEL_USER_ANTI = 5, 6

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_int64)


def build():
    """Compile the oracle library with its Makefile (gcc -O3 -fopenmp)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_nlevels.restype = ctypes.c_int
        L.or_nlevels.argtypes = [ctypes.c_int, ctypes.c_int]
        L.or_kernel_value.restype = ctypes.c_double
        L.or_kernel_value.argtypes = [ctypes.c_int, _dp, ctypes.c_double] + [ctypes.c_double] * 4
        L.or_kd_sort.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _lp, _ip, _ip]
        L.or_topology.argtypes = [ctypes.c_int, _ip, _ip, _ip]
        L.or_aca_points.restype = ctypes.c_int
        L.or_aca_points.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, _dp, ctypes.c_int, _dp, _ip, _ip]
        L.or_aca_matrix.restype = ctypes.c_int
        L.or_aca_matrix.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, _ip, _ip]
        L.or_solve_right.restype = ctypes.c_int
        L.or_solve_right.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_assemble.restype = ctypes.c_int
        L.or_assemble.argtypes = [ctypes.c_void_p]
        L.or_matvec.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.or_near_field.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.or_set_parallel_mode.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_set_exact_skeleton_rows.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_set_aca_wave_sums.argtypes = [ctypes.c_int]
        L.or_levels.restype = ctypes.c_int
        L.or_levels.argtypes = [ctypes.c_void_p]
        L.or_kernel_a.restype = ctypes.c_double
        L.or_kernel_a.argtypes = [ctypes.c_void_p]
        L.or_get_sorted_xy.argtypes = [ctypes.c_void_p, _dp]
        L.or_get_perm.argtypes = [ctypes.c_void_p, _lp]
        L.or_get_kd_counts.argtypes = [ctypes.c_void_p, _ip]
        L.or_get_kd_boxnum.argtypes = [ctypes.c_void_p, _ip]
        L.or_get_leaf_range.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _ip]
        L.or_get_rank.restype = ctypes.c_int
        L.or_get_rank.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.or_get_nin.restype = ctypes.c_int
        L.or_get_nin.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.or_get_skeleton.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip]
        L.or_get_l2p.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.or_get_m2l.restype = ctypes.c_int
        L.or_get_m2l.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.or_direct.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.or_direct_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, _dp]
        L.or_compute_error.restype = ctypes.c_double
        L.or_compute_error.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def nlevels(n, leaf):
    return lib().or_nlevels(int(n), int(leaf))


def kernel_value(kid, xi, yi, xj, yj, a=0.0, kparams=None):
    kp = np.zeros(4) if kparams is None else np.asarray(kparams, dtype=np.float64)
    return lib().or_kernel_value(kid, _d(kp), a, xi, yi, xj, yj)


def kd_sort(xy, L):
    """Returns (sorted_xy (n,2), perm (n,), kd_counts (4^L,), kd_boxnum (4^L,))."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    n = xy.shape[0]
    out = np.empty((n, 2), dtype=np.float64)
    perm = np.empty(n, dtype=np.int64)
    counts = np.zeros(4 ** L, dtype=np.int32)
    boxnum = np.zeros(4 ** L, dtype=np.int32)
    lib().or_kd_sort(n, _d(xy), L, _d(out), perm.ctypes.data_as(_lp), _i(counts), _i(boxnum))
    return out, perm, counts, boxnum


def topology(level):
    nb = 4 ** level
    nbr = np.empty((nb, 8), dtype=np.int32)
    inner = np.empty((nb, 16), dtype=np.int32)
    outer = np.empty((nb, 24), dtype=np.int32)
    lib().or_topology(level, _i(nbr), _i(inner), _i(outer))
    return nbr, inner, outer


def aca_points(kid, tol, rxy, cxy, a=0.0, kparams=None):
    kp = np.zeros(4) if kparams is None else np.asarray(kparams, dtype=np.float64)
    rxy = np.ascontiguousarray(rxy, dtype=np.float64)
    cxy = np.ascontiguousarray(cxy, dtype=np.float64)
    n1, n2 = rxy.shape[0], cxy.shape[0]
    mn = max(1, min(n1, n2))
    rp = np.zeros(mn, dtype=np.int32)
    cp = np.zeros(mn, dtype=np.int32)
    r = lib().or_aca_points(kid, _d(kp), a, tol, n1, _d(rxy), n2, _d(cxy), _i(rp), _i(cp))
    return r, rp[:r].copy(), cp[:r].copy()


def aca_matrix(A, tol):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n1, n2 = A.shape
    mn = max(1, min(n1, n2))
    rp = np.zeros(mn, dtype=np.int32)
    cp = np.zeros(mn, dtype=np.int32)
    r = lib().or_aca_matrix(n1, n2, _d(A), tol, _i(rp), _i(cp))
    return r, rp[:r].copy(), cp[:r].copy()


def solve_right(D, Ac):
    """X = Ac D^{-1} via LU of D^T (the L2P solve)."""
    D = np.ascontiguousarray(D, dtype=np.float64)
    Ac = np.ascontiguousarray(Ac, dtype=np.float64)
    r = D.shape[0]
    n1 = Ac.shape[0]
    X = np.empty((n1, r), dtype=np.float64)
    lib().or_solve_right(r, _d(D), n1, _d(Ac), _d(X))
    return X


class Tree:
    """The oracle operator: create -> assemble -> matvec (AFMM2D.hpp:14-33)."""

    def __init__(self, xy, leaf, tol_pow, kernel_id=KERNEL_LOG, kparams=None):
        xy = np.ascontiguousarray(xy, dtype=np.float64)
        self.n = xy.shape[0]
        kp = np.zeros(4) if kparams is None else np.asarray(kparams, dtype=np.float64)
        self._kp = kp
        self.h = lib().or_create(self.n, leaf, tol_pow, _d(xy), kernel_id, _d(kp))
        if not self.h:
            raise ValueError("oracle: invalid tree parameters (need nLevels >= 2)")
        self.L = lib().or_levels(self.h)

    def close(self):
        if self.h:
            lib().or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_parallel_mode(self, mode):
        lib().or_set_parallel_mode(self.h, mode)

    def set_exact_skeleton_rows(self, on=True):
        """Skeleton rows of L2P as the exact identity (the product's form)."""
        lib().or_set_exact_skeleton_rows(self.h, 1 if on else 0)

    def assemble(self):
        lib().or_assemble(self.h)

    def matvec(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        y = np.empty(self.n, dtype=np.float64)
        lib().or_matvec(self.h, _d(q), _d(y))
        return y

    def near_field(self, q):
        """The near-field term alone, caller order (AFMM2DTree.cpp:1199-1213)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        y = np.empty(self.n, dtype=np.float64)
        lib().or_near_field(self.h, _d(q), _d(y))
        return y

    def direct(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        y = np.empty(self.n, dtype=np.float64)
        lib().or_direct(self.h, _d(q), _d(y))
        return y

    def direct_rows(self, rows, q_sorted):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        q_sorted = np.ascontiguousarray(q_sorted, dtype=np.float64)
        out = np.empty(rows.shape[0], dtype=np.float64)
        lib().or_direct_rows(self.h, rows.shape[0], _i(rows), _d(q_sorted), _d(out))
        return out

    def compute_error(self, box):
        return lib().or_compute_error(self.h, box)

    @property
    def a(self):
        return lib().or_kernel_a(self.h)

    def sorted_xy(self):
        out = np.empty((self.n, 2), dtype=np.float64)
        lib().or_get_sorted_xy(self.h, _d(out))
        return out

    def perm(self):
        out = np.empty(self.n, dtype=np.int64)
        lib().or_get_perm(self.h, out.ctypes.data_as(_lp))
        return out

    def kd_counts(self):
        out = np.empty(4 ** self.L, dtype=np.int32)
        lib().or_get_kd_counts(self.h, _i(out))
        return out

    def kd_boxnum(self):
        out = np.empty(4 ** self.L, dtype=np.int32)
        lib().or_get_kd_boxnum(self.h, _i(out))
        return out

    def leaf_range(self, box):
        s = ctypes.c_int()
        c = ctypes.c_int()
        lib().or_get_leaf_range(self.h, box, ctypes.byref(s), ctypes.byref(c))
        return s.value, c.value

    def rank(self, level, box):
        return lib().or_get_rank(self.h, level, box)

    def ranks(self, level):
        return np.array([self.rank(level, k) for k in range(4 ** level)], dtype=np.int32)

    def nin(self, level, box):
        return lib().or_get_nin(self.h, level, box)

    def skeleton(self, level, box):
        r = self.rank(level, box)
        out = np.empty(max(r, 1), dtype=np.int32)
        lib().or_get_skeleton(self.h, level, box, _i(out))
        return out[:r]

    def l2p(self, level, box):
        r, n = self.rank(level, box), self.nin(level, box)
        out = np.empty(max(r * n, 1), dtype=np.float64)
        lib().or_get_l2p(self.h, level, box, _d(out))
        return out[: r * n].reshape(n, r)

    def m2l(self, level, box, slot, rank_il):
        r = self.rank(level, box)
        out = np.empty(max(r * rank_il, 1), dtype=np.float64)
        ok = lib().or_get_m2l(self.h, level, box, slot, _d(out))
        if not ok:
            return None
        return out[: r * rank_il].reshape(r, rank_il)


def set_aca_wave_sums(on):
    """The ACA's dot products in the engine's wave order (True) or in
    ascending order as the reference (False), for every tree."""
    lib().or_set_aca_wave_sums(1 if on else 0)
