"""DoubleBasis, symmetrizer and basis_embedding (LinearMap.jl:1-185) on the
device.

A DoubleBasis(B1, B2) describes maps from the source sector B2 to the target
sector B1: operator(mats, inds, DoubleBasis(B1, B2)) is the rectangular
operator (size(B1) x size(B2)) of the reference's column formula, and
T(v) is the basis-overlap map through the full tensor space.  Both run as
row gathers over B1 in libedkit_hip (edk_double_matvec, edk_double_map).
"""
import ctypes as C

import numpy as np

from ._lib import check, lib
from .device import DeviceArray, dev_ptr, dtype_code


class DoubleBasis:
    """DoubleBasis(B1, B2) (LinearMap.jl:88-113): B1 the target basis, B2 the
    source basis; check_doublebasis_compatible (:39-43) on construction."""

    def __init__(self, B1, B2):
        if B1.L != B2.L:
            raise ValueError("DoubleBasis requires bases with the same length.")
        if B1.base != B2.base:
            raise ValueError("DoubleBasis requires bases with the same local base.")
        self.B1, self.B2 = B1, B2
        self.L, self.base, self.ctx = B1.L, B1.base, B1.ctx
        self.complex_phases = bool(getattr(B1, "complex_phases", False) or getattr(B2, "complex_phases", False))

    @property
    def shape(self):
        return (self.B1.D, self.B2.D)

    def size(self, i=None):
        """size(b) / size(b, i) (LinearMap.jl:129-130)."""
        if i is None:
            return self.shape
        return self.shape[0] if i == 1 else self.shape[1] if i == 2 else 1

    def __len__(self):
        return self.L

    def __call__(self, v):
        """(B::DoubleBasis)(v) (LinearMap.jl:144-172): coordinates in B1 of
        the B2-coordinate vector v (or of each column of a (size(B2), nb) block)."""
        D1, D2 = self.shape
        host = isinstance(v, np.ndarray)
        shape = v.shape if hasattr(v, "shape") else (D2,)
        if shape[0] != D2:
            raise ValueError(f"Expected a vector of length {D2}; got {shape[0]}.")
        nb = 1 if len(shape) == 1 else int(np.prod(shape[1:]))
        if host:
            dt = np.complex64 if v.dtype == np.complex64 else np.complex128
            x = DeviceArray.from_host(np.ascontiguousarray(v, dt), ctx=self.ctx)
        else:
            x, dt = v, v.dtype
        out = DeviceArray((D1,) + tuple(shape[1:]), dt, ctx=self.ctx, zero=False)
        check(lib().edk_double_map(self.B1.h, self.B2.h, dtype_code(dt), dev_ptr(x), out.ptr, nb))
        return out.to_host() if host else out


def _columns(fn, D_in, D_out, chunk=64):
    """Dense D_out x D_in matrix of a linear map applied to identity blocks."""
    M = np.zeros((D_out, D_in), np.complex128)
    for c0 in range(0, D_in, chunk):
        c1 = min(D_in, c0 + chunk)
        E = np.zeros((D_in, c1 - c0), np.complex128)
        E[np.arange(c0, c1), np.arange(c1 - c0)] = 1
        M[:, c0:c1] = fn(E)
    return M


def symmetrizer(T):
    """symmetrizer(B::DoubleBasis) (LinearMap.jl:183-185): the dense
    size(B1) x size(B2) map embed(B1)^H embed(B2), i.e. T applied to every
    coordinate vector of B2."""
    D1, D2 = T.shape
    return _columns(T, D2, D1)


def basis_embedding(B):
    """basis_embedding(B) (LinearMap.jl:57-70): the dim(full) x size(B)
    embedding, entries coeff / orbit_order(B)."""
    from .basis import TensorBasis
    return symmetrizer(DoubleBasis(TensorBasis(L=B.L, base=B.base, ctx=B.ctx), B))


class DeviceDoubleOperator:
    """edk_dop handle of an Operator on a DoubleBasis."""

    def __init__(self, H):
        self.H = H
        T = H.B
        flat = H.flatten()
        self._flat = flat
        n, so, si, co, cp, no, rv, nz = flat
        h = C.c_void_p()
        check(lib().edk_double_operator_create(T.ctx.h, T.B1.h, T.B2.h, n, so.ctypes.data, si.ctypes.data,
                                               co.ctypes.data, cp.ctypes.data, no.ctypes.data, rv.ctypes.data,
                                               nz.ctypes.data, C.byref(h)))
        self.h = h

    def matvec(self, y, x, nb, accumulate):
        check(lib().edk_double_matvec(self.h, dtype_code(x.dtype), dev_ptr(y), dev_ptr(x), int(nb),
                                      int(accumulate)))
        return y

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().edk_double_operator_destroy(self.h)
            except Exception:
                pass
            self.h = None


def is_double(H):
    return isinstance(getattr(H, "B", None), DoubleBasis)


def double_operator(H):
    d = H._dev.get("dop")
    if d is None:
        d = DeviceDoubleOperator(H)
        H._dev["dop"] = d
    return d


def double_apply(H, v, target=None, accumulate=False):
    """H * v or mul!(target, H, v) for H on a DoubleBasis (rectangular)."""
    op = double_operator(H)
    D1, D2 = H.B.shape
    shape = v.shape if hasattr(v, "shape") else (D2,)
    if shape[0] != D2:
        raise ValueError(f"vector length {shape[0]} does not match the source basis dimension {D2}")
    nb = 1 if len(shape) == 1 else int(np.prod(shape[1:]))
    oshape = (D1,) + tuple(shape[1:])
    ctx = H.B.ctx
    if isinstance(v, np.ndarray):
        dt = np.complex64 if v.dtype == np.complex64 and (target is None or target.dtype == np.complex64) \
            else np.complex128
        x = DeviceArray.from_host(np.ascontiguousarray(v, dt), ctx=ctx)
        if accumulate:
            y = DeviceArray.from_host(np.ascontiguousarray(target, dt), ctx=ctx)
        else:
            y = DeviceArray(oshape, dt, ctx=ctx, zero=False)
        op.matvec(y, x, nb, accumulate)
        out = y.to_host()
        if target is not None:
            target[...] = out
            return target
        return out
    y = target if target is not None else DeviceArray(oshape, v.dtype, ctx=ctx, zero=False)
    op.matvec(y, v, nb, accumulate)
    return y


def double_array(H):
    """Array(opt) for H on a DoubleBasis: the dense size(B1) x size(B2) matrix."""
    D1, D2 = H.B.shape
    return _columns(lambda E: double_apply(H, E), D2, D1)
