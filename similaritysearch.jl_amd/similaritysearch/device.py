"""Device buffers, streams and events through libssg_hip.so (ssg_dev_*, ssg_stream_*,
ssg_event_*), so host code needs no other GPU runtime in the process."""
import ctypes as C

import numpy as np

from . import _lib


class DeviceArray:
    """A device allocation holding an array of `shape` and `dtype`."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._p = C.c_void_p()
        _lib.check(_lib.load().ssg_dev_malloc(C.byref(self._p), self.nbytes), "ssg_dev_malloc")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _lib.check(_lib.load().ssg_memcpy(d._p, a.ctypes.data, a.nbytes, 0), "ssg_memcpy")
        return d

    @property
    def ptr(self):
        return self._p.value

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _lib.check(_lib.load().ssg_memcpy(out.ctypes.data, self._p, self.nbytes, 1), "ssg_memcpy")
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        _lib.check(_lib.load().ssg_memcpy(self._p, a.ctypes.data, a.nbytes, 0), "ssg_memcpy")

    def free(self):
        if self._p:
            _lib.load().ssg_dev_free(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self._s = C.c_void_p()
        _lib.check(_lib.load().ssg_stream_create(C.byref(self._s)), "ssg_stream_create")

    @property
    def handle(self):
        return self._s

    def synchronize(self):
        _lib.check(_lib.load().ssg_stream_synchronize(self._s), "ssg_stream_synchronize")

    def __del__(self):
        try:
            _lib.load().ssg_stream_destroy(self._s)
        except Exception:
            pass


class Event:
    def __init__(self):
        self._e = C.c_void_p()
        _lib.check(_lib.load().ssg_event_create(C.byref(self._e)), "ssg_event_create")

    def record(self, stream):
        _lib.check(_lib.load().ssg_event_record(self._e, stream.handle), "ssg_event_record")

    def elapsed_ms(self, end):
        ms = C.c_float()
        _lib.check(_lib.load().ssg_event_elapsed_ms(self._e, end._e, C.byref(ms)), "elapsed")
        return ms.value

    def __del__(self):
        try:
            _lib.load().ssg_event_destroy(self._e)
        except Exception:
            pass


def synchronize():
    _lib.check(_lib.load().ssg_device_synchronize(), "ssg_device_synchronize")
