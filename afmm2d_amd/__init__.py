"""MI355X AFMM2D matvec engine: Python mirror of the reference's AFMM API.

The reference interface (include/AFMM2D.hpp:14-33) is

    AFMM(N, MinParticlesInLeaf, TOL_POW, loc)   # KD tree + quad tree
    afmm.assemble()                              # ACA skeletons, L2P, M2L
    y = afmm.computeMatVecProduct(b)             # y = K b, caller order
    afmm.evaluateError()                         # one random leaf vs direct sum

with the kernel supplied by ``userkernel::getMatrixEntry``.  Here the same
flow calls libafmm2d.so (hand-written HIP kernels for gfx950) through its C
ABI (include/afmm2d.h); the kernel is chosen from the built-in registry.
There is no CPU path: if the library is missing the import of a handle fails.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# AFMM2D_LIB names another build of the library (e.g. a tuned variant)
LIB_PATH = os.environ.get("AFMM2D_LIB") or os.path.join(_HERE, "libafmm2d.so")

KERNELS = {"log": 0, "inv_r": 1, "gaussian": 2, "reg_inv_r": 3, "user": 4, "user_sym": 5,
           "user_anti": 6, "user_symmetric": 7}

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_LEVELS = 2
ERR_STATE = 3
ERR_OUT_OF_MEMORY = 4
ERR_HIP = 5
ERR_UNSUPPORTED = 6
ERR_NOT_FOUND = 7

# Every symbol include/afmm2d.h declares.
EXPORTS = [
    "afmm2d_last_error", "afmm2d_status_string", "afmm2d_create", "afmm2d_destroy",
    "afmm2d_assemble", "afmm2d_matvec", "afmm2d_matvec_device", "afmm2d_levels",
    "afmm2d_get_permutation", "afmm2d_get_sorted_locations", "afmm2d_get_leaf_counts",
    "afmm2d_get_leaf_range", "afmm2d_get_interactions", "afmm2d_get_ranks",
    "afmm2d_get_skeleton", "afmm2d_get_l2p", "afmm2d_get_m2l", "afmm2d_near_field",
    "afmm2d_direct",
    "afmm2d_direct_rows", "afmm2d_compute_error", "afmm2d_evaluate_error",
    "afmm2d_get_matrix_entry", "afmm2d_kernel_value", "afmm2d_get_stats",
    "afmm2d_set_profiling", "afmm2d_get_kernel_time", "afmm2d_reset_kernel_times",
    "afmm2d_matvec_ex", "afmm2d_matvec_device_ex", "afmm2d_prepare_precision",
    "afmm2d_direct_rows_device_noalloc",
    "afmm2d_batch_assemble", "afmm2d_batch_create", "afmm2d_batch_matvec", "afmm2d_batch_tune",
    "afmm2d_batch_capture", "afmm2d_batch_get_info", "afmm2d_batch_destroy",
    "afmm2d_rank_model_create", "afmm2d_rank_model_destroy", "afmm2d_rank_model_get_params",
    "afmm2d_rank_model_set_params", "afmm2d_rank_model_predict", "afmm2d_rank_model_gradient",
    "afmm2d_rank_model_train", "afmm2d_set_rank_model", "afmm2d_get_rank_samples",
    "afmm2d_set_m2l_factoring", "afmm2d_get_m2l_storage", "afmm2d_check_symmetry",
    "afmm2d_set_m2l_storage",
    "afmm2d_release_scratch",
]

RANK_FEATURES = 16
RANK_PARAMS = 9473

# afmm2d_precision and the blocked-mode flag (include/afmm2d.h)
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16": 2}
MATVEC_BLOCKED = 0x100


class AFMM2DError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("afmm2d status %d: %s" % (status, msg))
        self.status = status


class Stats(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("levels", ctypes.c_int), ("tol_pow", ctypes.c_int),
        ("kernel_id", ctypes.c_int), ("max_rank", ctypes.c_int),
        ("skeleton_total", ctypes.c_int64), ("il_pairs", ctypes.c_int64),
        ("near_pairs", ctypes.c_int64), ("m2l_bytes", ctypes.c_int64),
        ("m2l_columns", ctypes.c_int64),
        ("l2p_bytes", ctypes.c_int64), ("matvec_bytes", ctypes.c_int64),
        ("create_seconds", ctypes.c_double), ("assemble_seconds", ctypes.c_double),
        ("m2l_leaf_bytes", ctypes.c_int64), ("m2l_leaf_columns", ctypes.c_int64),
        ("skeleton_leaf", ctypes.c_int64), ("near_waves", ctypes.c_int64),
        ("blocked_m2l_coeffs", ctypes.c_int64), ("blocked_basis_coeffs", ctypes.c_int64),
        ("rank_boxes", ctypes.c_int64), ("rank_fallbacks", ctypes.c_int64),
        ("aca_scratch_peak", ctypes.c_int64), ("aca_launches", ctypes.c_int64),
        ("aca_seconds", ctypes.c_double),
        ("m2l_factored_pairs", ctypes.c_int64), ("m2l_factored_bytes", ctypes.c_int64),
        ("m2l_factored_dense_bytes", ctypes.c_int64), ("blocked_m2l_coeffs_bf16", ctypes.c_int64),
        ("m2l_storage", ctypes.c_int64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)


class BatchInfo(ctypes.Structure):
    """afmm2d_batch_info (include/afmm2d.h)."""
    _fields_ = [(n, ctypes.c_int) for n in ("launches", "depth", "xcd", "near_waves", "rows_m2l",
                                             "captured", "tuned", "nproblems")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def library_path():
    return LIB_PATH


def lib():
    """Loads libafmm2d.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libafmm2d.so is not built (run __graft_entry__.build() or "
                          "python -m afmm2d_amd.build): the engine has no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    i, d, c = ctypes.c_int, ctypes.c_double, ctypes.c_char_p
    sig = {
        "afmm2d_last_error": (c, []),
        "afmm2d_status_string": (c, [i]),
        "afmm2d_create": (i, [i, i, i, _dp, i, i, _dp, i, ctypes.POINTER(_vp)]),
        "afmm2d_destroy": (i, [_vp]),
        "afmm2d_assemble": (i, [_vp]),
        "afmm2d_matvec": (i, [_vp, _dp, _dp, i]),
        "afmm2d_matvec_device": (i, [_vp, _vp, _vp, _vp]),
        "afmm2d_levels": (i, [_vp, _ip]),
        "afmm2d_get_permutation": (i, [_vp, _lp]),
        "afmm2d_get_sorted_locations": (i, [_vp, _dp]),
        "afmm2d_get_leaf_counts": (i, [_vp, _ip]),
        "afmm2d_get_leaf_range": (i, [_vp, i, _ip, _ip]),
        "afmm2d_get_interactions": (i, [_vp, i, _ip, _ip, _ip]),
        "afmm2d_get_ranks": (i, [_vp, i, _ip]),
        "afmm2d_get_skeleton": (i, [_vp, i, i, _ip, i, _ip]),
        "afmm2d_get_l2p": (i, [_vp, i, i, _dp, ctypes.c_int64, _ip, _ip]),
        "afmm2d_get_m2l": (i, [_vp, i, i, i, _dp, ctypes.c_int64, _ip, _ip]),
        "afmm2d_near_field": (i, [_vp, _dp, _dp]),
        "afmm2d_direct": (i, [_vp, _dp, _dp]),
        "afmm2d_direct_rows": (i, [_vp, i, _lp, _dp, _dp]),
        "afmm2d_compute_error": (i, [_vp, i, _dp]),
        "afmm2d_evaluate_error": (i, [_vp, ctypes.c_uint, _ip, _dp]),
        "afmm2d_get_matrix_entry": (i, [_vp, i, i, _dp]),
        "afmm2d_kernel_value": (i, [i, _dp, d, d, d, d, d, _dp]),
        "afmm2d_get_stats": (i, [_vp, ctypes.POINTER(Stats)]),
        "afmm2d_set_profiling": (i, [_vp, i]),
        "afmm2d_get_kernel_time": (i, [_vp, c, _dp, _lp]),
        "afmm2d_reset_kernel_times": (i, [_vp]),
        "afmm2d_matvec_ex": (i, [_vp, _dp, ctypes.c_int64, _dp, ctypes.c_int64, i, i]),
        "afmm2d_matvec_device_ex": (i, [_vp, _vp, ctypes.c_int64, _vp, ctypes.c_int64, i, i, _vp]),
        "afmm2d_prepare_precision": (i, [_vp, i]),
        "afmm2d_direct_rows_device_noalloc": (i, [_vp, i, _vp, _vp, _vp, _vp]),
        "afmm2d_batch_assemble": (i, [ctypes.POINTER(_vp), i]),
        "afmm2d_batch_create": (i, [ctypes.POINTER(_vp), i, ctypes.POINTER(_vp), _lp,
                                    ctypes.POINTER(_vp), _lp, i, i, ctypes.POINTER(_vp)]),
        "afmm2d_batch_matvec": (i, [_vp, _vp]),
        "afmm2d_batch_tune": (i, [_vp, _vp, _dp]),
        "afmm2d_batch_capture": (i, [_vp, i]),
        "afmm2d_batch_get_info": (i, [_vp, ctypes.POINTER(BatchInfo)]),
        "afmm2d_batch_destroy": (i, [_vp]),
        "afmm2d_rank_model_create": (i, [i, ctypes.c_uint64, ctypes.POINTER(_vp)]),
        "afmm2d_rank_model_destroy": (i, [_vp]),
        "afmm2d_rank_model_get_params": (i, [_vp, _fp]),
        "afmm2d_rank_model_set_params": (i, [_vp, _fp]),
        "afmm2d_rank_model_predict": (i, [_vp, _fp, ctypes.c_int64, _fp]),
        "afmm2d_rank_model_gradient": (i, [_vp, _fp, _fp, i, _fp, _dp]),
        "afmm2d_rank_model_train": (i, [_vp, _fp, _fp, ctypes.c_int64, i, i, ctypes.c_float,
                                        ctypes.c_uint64, _dp]),
        "afmm2d_set_rank_model": (i, [_vp, _vp, ctypes.c_float]),
        "afmm2d_get_rank_samples": (i, [_vp, _fp, _ip, ctypes.c_int64, _lp]),
        "afmm2d_set_m2l_factoring": (i, [_vp, i]),
        "afmm2d_set_m2l_storage": (i, [_vp, i]),
        "afmm2d_get_m2l_storage": (i, [_vp, i, i, i, _ip, _ip]),
        "afmm2d_check_symmetry": (i, [_vp, _ip]),
        "afmm2d_release_scratch": (i, [i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return _lib


def _check(status):
    if status != OK:
        msg = lib().afmm2d_last_error().decode(errors="replace")
        raise AFMM2DError(status, msg)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _kparams(kparams):
    """The 4 kernel parameters the C side reads, zero-padded."""
    kp = np.zeros(4, dtype=np.float64)
    if kparams is not None:
        v = np.asarray(kparams, dtype=np.float64).ravel()
        if v.size > 4:
            raise ValueError("at most 4 kernel parameters")
        kp[: v.size] = v
    return kp


def _mode(precision, blocked):
    if isinstance(precision, str):
        precision = PRECISIONS[precision]
    return int(precision) | (MATVEC_BLOCKED if blocked else 0)


def kernel_value(kernel, xi, yi, xj, yj, a=0.0, kparams=None):
    """The built-in kernel on explicit coordinates (host evaluation)."""
    kid = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
    kp = _kparams(kparams)
    out = ctypes.c_double()
    _check(lib().afmm2d_kernel_value(kid, _d(kp), a, xi, yi, xj, yj, ctypes.byref(out)))
    return out.value


def n_levels(N, MinParticlesInLeaf):
    """nLevels = (unsigned)(log(N/M)/log(4)) with integer N/M (AFMM2D.cpp:7)."""
    q = int(N) // int(MinParticlesInLeaf)
    return int(np.log(float(q)) / np.log(4.0)) if q > 0 else 0


class AFMM:
    """The operator: AFMM(N, MinParticlesInLeaf, TOL_POW, loc) (AFMM2D.hpp:14-33).

    loc: (N, 2) array of point coordinates.  kernel: "log", "inv_r",
    "gaussian" or "reg_inv_r" (the examples' a/R kernel).
    """

    def __init__(self, N, MinParticlesInLeaf, TOL_POW, loc, kernel="log", kparams=None,
                 device=-1):
        loc = np.ascontiguousarray(loc, dtype=np.float64)
        if loc.shape != (int(N), 2):
            raise ValueError("loc must have shape (N, 2)")
        self.N = int(N)
        self.MinParticlesInLeaf = int(MinParticlesInLeaf)
        self.TOL_POW = int(TOL_POW)
        self.kernel = kernel
        kid = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        kp = _kparams(kparams)
        h = _vp()
        _check(lib().afmm2d_create(self.N, self.MinParticlesInLeaf, self.TOL_POW, _d(loc), 0, kid,
                                   _d(kp), int(device), ctypes.byref(h)))
        self._h = h
        lv = ctypes.c_int()
        _check(lib().afmm2d_levels(self._h, ctypes.byref(lv)))
        self.nLevels = lv.value
        self._perm = None

    # lifecycle --------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().afmm2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # the reference flow -----------------------------------------------------
    def assemble(self):
        _check(lib().afmm2d_assemble(self._h))

    def computeMatVecProduct(self, b):
        """y = K b for b of shape (N,) or (N, R), in the caller's point order."""
        b = np.asarray(b, dtype=np.float64)
        one = b.ndim == 1
        B = np.asfortranarray(b.reshape(self.N, -1))
        Y = np.empty_like(B, order="F")
        _check(lib().afmm2d_matvec(self._h, _d(B), _d(Y), B.shape[1]))
        return Y[:, 0].copy() if one else Y

    def matvec_device(self, d_b, d_y, stream=None):
        """y = K b on device-resident vectors (device pointers as ints)."""
        _check(lib().afmm2d_matvec_device(self._h, _vp(d_b), _vp(d_y), _vp(stream or 0)))

    def matvec(self, b, precision="fp64", blocked=False):
        """Y = K B for B of shape (N,) or (N, R) in the caller's order, in the
        given precision ("fp64", "fp32", "bf16"); blocked=True runs the
        several-RHS passes also for one FP64 column (afmm2d_matvec_ex)."""
        b = np.asarray(b, dtype=np.float64)
        one = b.ndim == 1
        B = np.asfortranarray(b.reshape(self.N, -1))
        Y = np.empty_like(B, order="F")
        _check(lib().afmm2d_matvec_ex(self._h, _d(B), self.N, _d(Y), self.N, B.shape[1],
                                      _mode(precision, blocked)))
        return Y[:, 0].copy() if one else Y

    def matvec_device_ex(self, d_b, ldb, d_y, ldy, nrhs, precision="fp64", blocked=False,
                         stream=None):
        """Y = K B on device-resident column-major arrays (device pointers as
        ints), afmm2d_matvec_device_ex."""
        _check(lib().afmm2d_matvec_device_ex(self._h, _vp(d_b), int(ldb), _vp(d_y), int(ldy),
                                             int(nrhs), _mode(precision, blocked),
                                             _vp(stream or 0)))

    def prepare_precision(self, precision):
        """Builds the factors of a precision now (afmm2d_prepare_precision)."""
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        _check(lib().afmm2d_prepare_precision(self._h, int(precision)))

    def evaluateError(self, seed=0):
        """One leaf's relative error vs the direct sum (AFMM2D.cpp:67-74)."""
        box = ctypes.c_int()
        err = ctypes.c_double()
        _check(lib().afmm2d_evaluate_error(self._h, seed, ctypes.byref(box), ctypes.byref(err)))
        return box.value, err.value

    def compute_error(self, box):
        err = ctypes.c_double()
        _check(lib().afmm2d_compute_error(self._h, int(box), ctypes.byref(err)))
        return err.value

    # tree data ----------------------------------------------------------------
    @property
    def sorted_Properties(self):
        """perm[i] = caller index of KD-sorted point i."""
        if self._perm is None:
            p = np.empty(self.N, dtype=np.int64)
            _check(lib().afmm2d_get_permutation(self._h, p.ctypes.data_as(_lp)))
            self._perm = p
        return self._perm

    @property
    def sorted_Locations(self):
        xy = np.empty((self.N, 2), dtype=np.float64)
        _check(lib().afmm2d_get_sorted_locations(self._h, _d(xy)))
        return xy

    @property
    def NumberOfParticlesInLeaves(self):
        c = np.empty(4 ** self.nLevels, dtype=np.int32)
        _check(lib().afmm2d_get_leaf_counts(self._h, _i(c)))
        return c

    def leaf_range(self, box):
        s, c = ctypes.c_int(), ctypes.c_int()
        _check(lib().afmm2d_get_leaf_range(self._h, int(box), ctypes.byref(s), ctypes.byref(c)))
        return s.value, c.value

    def interactions(self, level):
        nb = 4 ** level
        nbr = np.empty((nb, 8), dtype=np.int32)
        inner = np.empty((nb, 16), dtype=np.int32)
        outer = np.empty((nb, 24), dtype=np.int32)
        _check(lib().afmm2d_get_interactions(self._h, level, _i(nbr), _i(inner), _i(outer)))
        return nbr, inner, outer

    def ranks(self, level):
        r = np.empty(4 ** level, dtype=np.int32)
        _check(lib().afmm2d_get_ranks(self._h, level, _i(r)))
        return r

    def skeleton(self, level, box):
        cap = 4096
        idx = np.empty(cap, dtype=np.int32)
        cnt = ctypes.c_int()
        _check(lib().afmm2d_get_skeleton(self._h, level, box, _i(idx), cap, ctypes.byref(cnt)))
        return idx[: cnt.value].copy()

    def l2p(self, level, box):
        rows, cols = ctypes.c_int(), ctypes.c_int()
        cap = 1 << 20
        out = np.empty(cap, dtype=np.float64)
        _check(lib().afmm2d_get_l2p(self._h, level, box, _d(out), cap, ctypes.byref(rows),
                                    ctypes.byref(cols)))
        return out[: rows.value * cols.value].reshape(rows.value, cols.value).copy()

    def m2l(self, level, box, slot):
        rows, cols = ctypes.c_int(), ctypes.c_int()
        cap = 1 << 20
        out = np.empty(cap, dtype=np.float64)
        st = lib().afmm2d_get_m2l(self._h, level, box, slot, _d(out), cap, ctypes.byref(rows),
                                  ctypes.byref(cols))
        if st == ERR_NOT_FOUND:
            return None
        _check(st)
        return out[: rows.value * cols.value].reshape(rows.value, cols.value).copy()

    def m2l_storage(self, level, box, slot):
        """(factor rank or -1 for a dense block, whether the stored block is
        K(skeleton_box, skeleton_slot)), or None for an empty slot."""
        p, first = ctypes.c_int(), ctypes.c_int()
        st = lib().afmm2d_get_m2l_storage(self._h, level, box, slot, ctypes.byref(p),
                                          ctypes.byref(first))
        if st == ERR_NOT_FOUND:
            return None
        _check(st)
        return p.value, bool(first.value)

    def getMatrixEntry(self, i, j):
        """kernel::getMatrixEntry on KD-sorted indices (kernel.hpp:23-26)."""
        v = ctypes.c_double()
        _check(lib().afmm2d_get_matrix_entry(self._h, int(i), int(j), ctypes.byref(v)))
        return v.value

    # checks and measurement -------------------------------------------------
    def near_field(self, b):
        """The near-field term alone (leaf self and neighbour blocks) on the
        device, caller order (AFMM2DTree.cpp:1199-1213)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        y = np.empty(self.N, dtype=np.float64)
        _check(lib().afmm2d_near_field(self._h, _d(b), _d(y)))
        return y

    def direct(self, b):
        """Dense direct sum y = K b on the device (tutorial.cpp:97-131)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        y = np.empty(self.N, dtype=np.float64)
        _check(lib().afmm2d_direct(self._h, _d(b), _d(y)))
        return y

    def direct_rows(self, rows, b):
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.empty(rows.shape[0], dtype=np.float64)
        _check(lib().afmm2d_direct_rows(self._h, rows.shape[0], rows.ctypes.data_as(_lp), _d(b),
                                        _d(out)))
        return out

    def check_symmetry(self):
        """False when K(i, j) != K(j, i) on a sample of this point set's pairs."""
        v = ctypes.c_int()
        _check(lib().afmm2d_check_symmetry(self._h, ctypes.byref(v)))
        return bool(v.value)

    def set_m2l_factoring(self, on=True):
        """Store compressible M2L blocks as ACA factors (before assemble)."""
        _check(lib().afmm2d_set_m2l_factoring(self._h, 1 if on else 0))

    def set_m2l_storage(self, storage="fp64"):
        """Dense M2L block storage (before assemble): "fp64", or "bfp48"
        (48-bit block floating point: 6 bytes per entry, each within 2^-46
        of its block's largest entry)."""
        codes = {"fp64": 0, "bfp48": 1}
        if storage not in codes:
            raise ValueError("M2L storage must be one of %s" % sorted(codes))
        _check(lib().afmm2d_set_m2l_storage(self._h, codes[storage]))

    def set_rank_model(self, model, margin=0.5):
        """Assemble with ACA capacities from model (None: the plain ACA)."""
        _check(lib().afmm2d_set_rank_model(self._h, model._m if model is not None else None,
                                           float(margin)))
        self._rank_model = model  # the model outlives the assembly

    def rank_samples(self):
        """(features (n, 16) float32, ranks (n,) int32): one per box of
        levels L..2, from the assembly."""
        cnt = ctypes.c_int64()
        _check(lib().afmm2d_get_rank_samples(self._h, None, None, 0, ctypes.byref(cnt)))
        X = np.zeros((cnt.value, RANK_FEATURES), dtype=np.float32)
        r = np.zeros(cnt.value, dtype=np.int32)
        _check(lib().afmm2d_get_rank_samples(self._h, X.ctypes.data_as(_fp), _i(r), cnt.value,
                                             ctypes.byref(cnt)))
        return X, r

    def stats(self):
        st = Stats()
        _check(lib().afmm2d_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def set_profiling(self, on, m2l_only=False):
        """Phase timing with HIP events (see afmm2d_set_profiling); with
        m2l_only only the leaf-level M2L launch is timed."""
        _check(lib().afmm2d_set_profiling(self._h, (2 if m2l_only else 1) if on else 0))

    def kernel_time(self, name):
        ms, n = ctypes.c_double(), ctypes.c_int64()
        _check(lib().afmm2d_get_kernel_time(self._h, name.encode(), ctypes.byref(ms),
                                            ctypes.byref(n)))
        return ms.value, n.value

    def reset_kernel_times(self):
        _check(lib().afmm2d_reset_kernel_times(self._h))


def release_scratch(device=-1):
    """Frees the ACA scratch the library keeps between assemblies."""
    _check(lib().afmm2d_release_scratch(int(device)))


def batch_assemble(problems):
    """Assembles every AFMM of the list (afmm2d_batch_assemble)."""
    hs = (_vp * len(problems))(*[p._h for p in problems])
    _check(lib().afmm2d_batch_assemble(hs, len(problems)))


class Batch:
    """Several problems' blocked matvecs as one (afmm2d_batch_*).

    problems: assembled AFMM objects; d_b, d_y: device pointers (ints) of
    each problem's charges and outputs, column-major with leading dimensions
    ldb, ldy (default N of each problem); nrhs columns (<= 32)."""

    def __init__(self, problems, d_b, d_y, nrhs, precision="fp64", ldb=None, ldy=None):
        n = len(problems)
        self._problems = list(problems)  # the handles outlive the batch
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        ldb = [p.N for p in problems] if ldb is None else list(ldb)
        ldy = [p.N for p in problems] if ldy is None else list(ldy)
        hs = (_vp * n)(*[p._h for p in problems])
        bs = (_vp * n)(*d_b)
        ys = (_vp * n)(*d_y)
        lb = np.asarray(ldb, dtype=np.int64)
        ly = np.asarray(ldy, dtype=np.int64)
        h = _vp()
        _check(lib().afmm2d_batch_create(hs, n, bs, lb.ctypes.data_as(_lp), ys,
                                         ly.ctypes.data_as(_lp), int(nrhs), int(precision),
                                         ctypes.byref(h)))
        self._b = h

    def close(self):
        if getattr(self, "_b", None) and self._b.value:
            lib().afmm2d_batch_destroy(self._b)
            self._b = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def matvec(self, stream=None):
        _check(lib().afmm2d_batch_matvec(self._b, _vp(stream or 0)))

    def tune(self, stream=None):
        ms = ctypes.c_double()
        _check(lib().afmm2d_batch_tune(self._b, _vp(stream or 0), ctypes.byref(ms)))
        return ms.value

    def capture(self, on=True):
        _check(lib().afmm2d_batch_capture(self._b, 1 if on else 0))

    def info(self):
        inf = BatchInfo()
        _check(lib().afmm2d_batch_get_info(self._b, ctypes.byref(inf)))
        return inf.as_dict()


class RankModel:
    """The learned rank predictor (afmm2d_rank_model_*): a 16-64-64-64-1
    FP32 MLP trained and evaluated by the library's kernels on the GPU."""

    def __init__(self, seed=0, device=-1):
        h = _vp()
        _check(lib().afmm2d_rank_model_create(int(device), int(seed), ctypes.byref(h)))
        self._m = h

    def close(self):
        if getattr(self, "_m", None) and self._m.value:
            lib().afmm2d_rank_model_destroy(self._m)
            self._m = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(RANK_PARAMS, dtype=np.float32)
        _check(lib().afmm2d_rank_model_get_params(self._m, p.ctypes.data_as(_fp)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == RANK_PARAMS
        _check(lib().afmm2d_rank_model_set_params(self._m, p.ctypes.data_as(_fp)))

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, RANK_FEATURES)
        out = np.zeros(len(X), dtype=np.float32)
        _check(lib().afmm2d_rank_model_predict(self._m, X.ctypes.data_as(_fp), len(X),
                                               out.ctypes.data_as(_fp)))
        return out

    def gradient(self, X, y):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, RANK_FEATURES)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(RANK_PARAMS, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().afmm2d_rank_model_gradient(self._m, X.ctypes.data_as(_fp),
                                                y.ctypes.data_as(_fp), len(X),
                                                g.ctypes.data_as(_fp), ctypes.byref(loss)))
        return loss.value, g

    def train(self, X, y, epochs=10, batch=1024, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, RANK_FEATURES)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().afmm2d_rank_model_train(self._m, X.ctypes.data_as(_fp),
                                             y.ctypes.data_as(_fp), len(X), int(epochs),
                                             int(batch), float(lr), int(seed),
                                             ctypes.byref(loss)))
        return loss.value


def rank_target(ranks):
    """The rank model's training target for ACA ranks: log2(1 + rank)."""
    return np.log2(1.0 + np.asarray(ranks, dtype=np.float64)).astype(np.float32)
