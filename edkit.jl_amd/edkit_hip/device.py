"""Device context and device arrays over the C ABI (no torch dependency).

A DeviceArray is a flat block of HBM owned by the library's allocator; any
object with a `data_ptr()` method (a torch tensor on the same device) can be
passed wherever a device vector is expected.
"""
import ctypes as C
import threading

import numpy as np

from ._lib import check, lib

_tls = threading.local()


class Context:
    """edk_ctx: one device, one HIP stream (NULL stream arg = private stream)."""

    def __init__(self, device=0, stream=None):
        h = C.c_void_p()
        check(lib().edk_ctx_create(device, stream, C.byref(h)))
        self.h = h
        self.device = device

    def synchronize(self):
        check(lib().edk_ctx_synchronize(self.h))

    @property
    def stream(self):
        return lib().edk_ctx_stream(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().edk_ctx_destroy(self.h)
            self.h = None


def default_context():
    """Per-thread default context on device 0 (contexts are not reentrant)."""
    ctx = getattr(_tls, "ctx", None)
    if ctx is None:
        ctx = Context(0)
        _tls.ctx = ctx
    return ctx


def set_default_context(ctx):
    _tls.ctx = ctx


_DT = {np.dtype(np.complex128): 0, np.dtype(np.complex64): 1}


class DeviceArray:
    """Flat device buffer with a numpy dtype and shape."""

    def __init__(self, shape, dtype=np.complex128, ctx=None, zero=True):
        self.ctx = ctx or default_context()
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().edk_malloc(self.ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value
        if zero:
            check(lib().edk_memset(self.ctx.h, self.ptr, 0, self.nbytes))

    @classmethod
    def from_host(cls, a, dtype=None, ctx=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.shape, a.dtype, ctx, zero=False)
        check(lib().edk_memcpy_htod(d.ctx.h, d.ptr, a.ctypes.data, d.nbytes))
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().edk_memcpy_dtoh(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().edk_memcpy_htod(self.ctx.h, self.ptr, a.ctypes.data, self.nbytes))

    def zero_(self):
        check(lib().edk_memset(self.ctx.h, self.ptr, 0, self.nbytes))
        return self

    def data_ptr(self):
        return self.ptr

    def __del__(self):
        if getattr(self, "ptr", None):
            try:
                lib().edk_free(self.ctx.h, self.ptr)
            except Exception:
                pass
            self.ptr = None


def dev_ptr(v):
    """Device pointer of a DeviceArray or a torch tensor."""
    if isinstance(v, DeviceArray):
        return v.ptr
    if hasattr(v, "data_ptr"):
        return v.data_ptr()
    raise TypeError("expected a device array (DeviceArray or a tensor on the device)")


def dtype_code(dt):
    dt = np.dtype(dt) if not hasattr(dt, "is_complex") else dt
    if dt in _DT:
        return _DT[dt]
    s = str(dt)
    if "complex128" in s:
        return 0
    if "complex64" in s:
        return 1
    raise TypeError(f"vectors must be complex128 or complex64, got {dt}")
