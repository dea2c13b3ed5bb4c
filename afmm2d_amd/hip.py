"""Device buffers, streams and events for device-resident use of libafmm2d.so
from Python, through the HIP runtime the library itself links
(libamdhip64.so.7 of /opt/rocm).

A process can hold another HIP runtime too (PyTorch ships its own): device
pointers are valid across runtimes, but streams and events are not, so work
is timed and ordered here with this runtime's objects, on the streams the
library's kernels run on.
"""
import ctypes

import numpy as np

_rt = None
_vp = ctypes.c_void_p


def runtime():
    """The HIP runtime of libafmm2d.so (loaded by the library itself)."""
    global _rt
    if _rt is not None:
        return _rt
    from . import lib
    lib()  # loads libamdhip64.so.7 as a dependency of libafmm2d.so
    last = None
    for name in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so.7", "libamdhip64.so"):
        try:
            h = ctypes.CDLL(name)
            break
        except OSError as e:
            last = e
    else:
        raise OSError("HIP runtime not found: %s" % last)
    h.hipMalloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
    h.hipFree.argtypes = [_vp]
    h.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
    h.hipMemset.argtypes = [_vp, ctypes.c_int, ctypes.c_size_t]
    h.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(_vp), ctypes.c_uint]
    h.hipStreamSynchronize.argtypes = [_vp]
    h.hipStreamDestroy.argtypes = [_vp]
    h.hipEventCreate.argtypes = [ctypes.POINTER(_vp)]
    h.hipEventRecord.argtypes = [_vp, _vp]
    h.hipEventSynchronize.argtypes = [_vp]
    h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), _vp, _vp]
    h.hipEventDestroy.argtypes = [_vp]
    h.hipDeviceSynchronize.argtypes = []
    h.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
    _rt = h
    return _rt


def _check(st, what):
    if st != 0:
        raise RuntimeError("%s failed with HIP status %d" % (what, st))


def device_count():
    n = ctypes.c_int(0)
    st = runtime().hipGetDeviceCount(ctypes.byref(n))
    return n.value if st == 0 else 0


def synchronize():
    _check(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A device allocation holding a numpy array's bytes."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.p = _vp()
        _check(runtime().hipMalloc(ctypes.byref(self.p), max(self.nbytes, 8)), "hipMalloc")

    @classmethod
    def from_numpy(cls, a):
        """A device copy of a's buffer in its own memory order (C or Fortran)."""
        if not (a.flags.c_contiguous or a.flags.f_contiguous):
            a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    @property
    def ptr(self):
        return self.p.value

    def upload(self, a):
        if not (a.flags.c_contiguous or a.flags.f_contiguous):
            a = np.ascontiguousarray(a)
        if a.nbytes:
            _check(runtime().hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")

    def download(self, out):
        if out.nbytes:
            _check(runtime().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def copy_from(self, other, nbytes=None):
        n = self.nbytes if nbytes is None else int(nbytes)
        _check(runtime().hipMemcpy(self.p, other.p, n, 3), "hipMemcpy D2D")

    def fill_zero(self):
        _check(runtime().hipMemset(self.p, 0, self.nbytes), "hipMemset")

    def free(self):
        if self.p:
            runtime().hipFree(self.p)
            self.p = _vp()


class Stream:
    """A non-blocking stream (default priority)."""

    def __init__(self):
        self.s = _vp()
        _check(runtime().hipStreamCreateWithFlags(ctypes.byref(self.s), 1), "hipStreamCreate")

    @property
    def handle(self):
        return self.s.value

    def synchronize(self):
        _check(runtime().hipStreamSynchronize(self.s), "hipStreamSynchronize")

    def destroy(self):
        if self.s:
            runtime().hipStreamDestroy(self.s)
            self.s = _vp()


class Event:
    def __init__(self):
        self.e = _vp()
        _check(runtime().hipEventCreate(ctypes.byref(self.e)), "hipEventCreate")

    def record(self, stream=None):
        _check(runtime().hipEventRecord(self.e, _vp(stream or 0)), "hipEventRecord")

    def synchronize(self):
        _check(runtime().hipEventSynchronize(self.e), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _check(runtime().hipEventElapsedTime(ctypes.byref(ms), self.e, end.e), "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        if self.e:
            runtime().hipEventDestroy(self.e)
            self.e = _vp()
