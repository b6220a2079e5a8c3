"""Schmidt decomposition and entanglement entropy of sector states
(Schmidt.jl:100-330, AbelianBasis.jl:704-719): the Schmidt matrix is built on
the device by orbit expansion, the squared Schmidt values come from the
device eigenvalues of the reduced density matrix."""
import ctypes as C

import numpy as np

from ._lib import check, lib
from .device import DeviceArray, dev_ptr, dtype_code


def _prepare(v, Ainds, B):
    x = v if isinstance(v, DeviceArray) else DeviceArray.from_host(
        np.asarray(v, np.complex64 if np.asarray(v).dtype == np.complex64 else np.complex128), ctx=B.ctx)
    if x.shape[0] != B.D:
        raise ValueError("state length does not match the basis dimension")
    A = np.ascontiguousarray(np.asarray(list(Ainds), np.int32))
    return x, A


def schmidt(v, Ainds, B):
    """schmidt(v, Ainds, b): the base^|A| x base^(L-|A|) Schmidt matrix (host)."""
    x, A = _prepare(v, Ainds, B)
    dA, dB = B.base ** len(A), B.base ** (B.L - len(A))
    M = DeviceArray((dA * dB,), np.complex128, ctx=B.ctx, zero=False)
    check(lib().edk_schmidt(B.h, dtype_code(x.dtype), dev_ptr(x), len(A),
                            A.ctypes.data if len(A) else None, M.ptr))
    return M.to_host().reshape((dB, dA)).T


def ent_spec(v, Ainds, B):
    """ent_spec: the Schmidt values (svdvals(schmidt(v, Ainds, b))), descending."""
    return np.sqrt(_squared_spec(v, Ainds, B))


def _squared_spec(v, Ainds, B):
    x, A = _prepare(v, Ainds, B)
    n = min(B.base ** len(A), B.base ** (B.L - len(A)))
    s2 = np.zeros(n, np.float64)
    cnt = C.c_int64()
    check(lib().edk_ent_spec(B.h, dtype_code(x.dtype), dev_ptr(x), len(A),
                             A.ctypes.data if len(A) else None, s2.ctypes.data, C.byref(cnt)))
    return s2[:cnt.value]


def entropy(s, alpha=1, cutoff=1e-20):
    """entropy(s; α, cutoff) (Schmidt.jl:100-140) of a descending probability vector:
    Shannon for α = 1 and the support size for α = 0 (both stop at the first
    value not above cutoff), log(sum(s^α))/(1-α) otherwise."""
    s = np.asarray(s, np.float64)
    if alpha == 1:
        ent = 0.0
        for si in s:
            if not si > cutoff:
                break
            ent -= si * np.log(si)
        return ent
    if alpha == 0:
        n = 0
        for si in s:
            if not si > cutoff:
                break
            n += 1
        return n
    return float(np.log(np.sum(s ** alpha)) / (1 - alpha))


def ent_S(v, Ainds, B, alpha=1, cutoff=1e-20):
    """ent_S(v, Aind, b; α, cutoff) (Schmidt.jl:150-154): the entropy of the
    squared Schmidt values."""
    return entropy(_squared_spec(v, Ainds, B), alpha=alpha, cutoff=cutoff)
