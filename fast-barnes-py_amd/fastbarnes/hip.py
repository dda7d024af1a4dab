"""
Minimal ctypes access to the HIP runtime (libamdhip64): device buffers, streams and
events for callers that do not use PyTorch.  Only plumbing: every computation runs
in libfastbarnes.so.
"""
import ctypes

import numpy as np

_hip = None

H2D = 1
D2H = 2
D2D = 3


def _load():
    global _hip
    if _hip is None:
        # the runtime instance libfastbarnes.so is bound to (resolved by soname)
        from fastbarnes import _lib
        _lib.lib()
        try:
            _hip = ctypes.CDLL('libamdhip64.so.7')
        except OSError as e:
            raise RuntimeError('cannot load the HIP runtime: %s' % e)
        vp = ctypes.c_void_p
        _hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        _hip.hipFree.argtypes = [vp]
        _hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpyAsync.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
        _hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(vp), ctypes.c_uint]
        _hip.hipStreamSynchronize.argtypes = [vp]
        _hip.hipStreamDestroy.argtypes = [vp]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        _hip.hipEventRecord.argtypes = [vp, vp]
        _hip.hipEventSynchronize.argtypes = [vp]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        _hip.hipEventDestroy.argtypes = [vp]
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipGetErrorString.argtypes = [ctypes.c_int]
        _hip.hipSetDevice.argtypes = [ctypes.c_int]
    return _hip


def check(rc):
    if rc != 0:
        raise RuntimeError('HIP error %d: %s' % (rc, _load().hipGetErrorString(rc).decode()))


def set_device(i):
    check(_load().hipSetDevice(i))


def synchronize():
    check(_load().hipDeviceSynchronize())


class Stream:
    """A non-blocking HIP stream of default priority."""

    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(_load().hipStreamCreateWithFlags(ctypes.byref(self.handle), 1))

    @property
    def ptr(self):
        return self.handle.value

    def synchronize(self):
        check(_load().hipStreamSynchronize(self.handle))

    def __del__(self):
        if _hip is not None and self.handle:
            _hip.hipStreamDestroy(self.handle)
            self.handle = ctypes.c_void_p()


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(_load().hipEventCreate(ctypes.byref(self.handle)))

    def record(self, stream=None):
        check(_load().hipEventRecord(self.handle, stream.handle if stream is not None else None))

    def synchronize(self):
        check(_load().hipEventSynchronize(self.handle))

    def elapsed_ms(self, end):
        t = ctypes.c_float()
        check(_load().hipEventElapsedTime(ctypes.byref(t), self.handle, end.handle))
        return t.value

    def __del__(self):
        if _hip is not None and self.handle:
            _hip.hipEventDestroy(self.handle)
            self.handle = ctypes.c_void_p()


class DeviceArray:
    """A device buffer with a numpy dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.handle = ctypes.c_void_p()
        check(_load().hipMalloc(ctypes.byref(self.handle), max(self.nbytes, 1)))

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @property
    def ptr(self):
        return self.handle.value

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError('size mismatch')
        check(_load().hipMemcpy(self.handle, a.ctypes.data, self.nbytes, H2D))

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(_load().hipMemcpy(out.ctypes.data, self.handle, self.nbytes, D2H))
        return out

    def free(self):
        if _hip is not None and self.handle:
            _hip.hipFree(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        self.free()
