"""Device buffers, streams and events on top of libdcc (no torch)."""
import ctypes

import numpy as np

from distr_cc import _lib as L

_DEVICE_READY = {}


def device_count():
    n = ctypes.c_int(0)
    L.call("dcc_device_count", ctypes.byref(n))
    return n.value


def mem_get_info():
    """(free, total) device memory in bytes."""
    f, t = ctypes.c_int64(), ctypes.c_int64()
    L.call("dcc_mem_get_info", ctypes.byref(f), ctypes.byref(t))
    return int(f.value), int(t.value)


def init(device=0):
    """Select ``device`` for this thread; returns its gfx arch name."""
    if device in _DEVICE_READY:
        return _DEVICE_READY[device]
    if device_count() <= device:
        raise RuntimeError("no HIP device %d visible (libdcc needs an MI355X / gfx950 GPU)" % device)
    buf = ctypes.create_string_buffer(64)
    L.call("dcc_set_device", device, buf, 64)
    arch = buf.value.decode()
    if not arch.startswith("gfx950"):
        raise RuntimeError("libdcc is built for gfx950 only; device %d is %s" % (device, arch))
    _DEVICE_READY[device] = arch
    return arch


class Stream:
    def __init__(self, handle=None):
        if handle is None:
            h = ctypes.c_void_p()
            L.call("dcc_stream_create", ctypes.byref(h))
            self.handle, self._own = h, True
        else:
            self.handle, self._own = ctypes.c_void_p(handle), False

    def synchronize(self):
        L.call("dcc_stream_synchronize", self.handle)

    def close(self):
        if self._own and self.handle is not None and self.handle.value:
            L.call("dcc_stream_destroy", self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        L.call("dcc_event_create", ctypes.byref(self.handle))

    def record(self, stream=None):
        L.call("dcc_event_record", self.handle, stream.handle if stream is not None else None)

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        L.call("dcc_event_elapsed_ms", ctypes.byref(ms), self.handle, later.handle)
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle is not None and self.handle.value:
                L.call("dcc_event_destroy", self.handle)
        except Exception:
            pass


class DeviceArray:
    """A C-contiguous device array (owning unless created with ``owner=False``)."""

    def __init__(self, shape, dtype=np.float64, ptr=None, owner=True):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        if ptr is None:
            p = ctypes.c_void_p()
            L.call("dcc_malloc", ctypes.byref(p), self.nbytes)
            self.ptr, self._owner = p.value or 0, True
        else:
            self.ptr, self._owner = int(ptr), bool(owner)

    @property
    def vp(self):
        return ctypes.c_void_p(self.ptr)

    def offset_ptr(self, elements):
        return ctypes.c_void_p(self.ptr + int(elements) * self.dtype.itemsize)

    @classmethod
    def from_host(cls, arr, stream=None, dtype=None):
        a = np.ascontiguousarray(arr, dtype=dtype if dtype is not None else np.asarray(arr).dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    @classmethod
    def zeros(cls, shape, dtype=np.float64, stream=None):
        d = cls(shape, dtype)
        if d.nbytes:
            if d.dtype in (np.float64, np.float32):
                d.fill(0.0, stream)
            else:
                d.upload(np.zeros(d.shape, d.dtype), stream)
        return d

    def upload(self, arr, stream=None):
        a = L._as_array(arr, self.dtype, "host array")
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch: host %d bytes, device %d bytes" % (a.nbytes, self.nbytes))
        L.call("dcc_memcpy_htod", self.vp, a.ctypes.data_as(ctypes.c_void_p), self.nbytes,
               stream.handle if stream else None)

    def download(self, out=None, stream=None):
        if out is None:
            out = np.empty(self.shape, self.dtype)
        out = L._as_array(out, self.dtype, "host array")
        L.call("dcc_memcpy_dtoh", out.ctypes.data_as(ctypes.c_void_p), self.vp, self.nbytes,
               stream.handle if stream else None)
        return out

    def fill(self, value, stream=None):
        fn = {np.dtype(np.float64): "dcc_fill_f64", np.dtype(np.float32): "dcc_fill_f32"}[self.dtype]
        L.call(fn, self.vp, self.size, float(value), stream.handle if stream else None)

    def free(self):
        if self._owner and self.ptr:
            L.call("dcc_free", ctypes.c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
