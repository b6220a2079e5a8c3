"""mul!(y, H, x), H*x and mul(H, x) on the device (Operator.jl:309-394).

Host numpy vectors go through edk_matvec_host (copy in, gather kernel, copy
out); device vectors (DeviceArray or a tensor on the device) go straight to
edk_matvec.  ComplexF64 vectors run the FP64 kernel, ComplexF32 vectors the
FP32 kernel.  A 2-D (D, nb) array is a block of nb right-hand sides.
"""
import collections
import ctypes as C

import numpy as np

from ._lib import check, lib
from .device import DeviceArray, dev_ptr, dtype_code
from .linearmap import double_apply, double_array, is_double


class DeviceOperator:
    """edk_op handle of an Operator on its basis' device."""

    def __init__(self, H):
        self.H = H
        self.B = H.B
        flat = H.flatten()
        self._flat = flat
        h = C.c_void_p()
        n, so, si, co, cp, no, rv, nz = flat
        check(lib().edk_operator_create(self.B.ctx.h, self.B.h, n, so.ctypes.data, si.ctypes.data,
                                        co.ctypes.data, cp.ctypes.data, no.ctypes.data,
                                        rv.ctypes.data, nz.ctypes.data, C.byref(h)))
        self.h = h

    def set_variant(self, dtype, nb, variant):
        """Force a schedule (EDK_VARIANT_*) for this dtype and block width."""
        check(lib().edk_operator_set_variant(self.h, dtype_code(np.dtype(dtype)), int(nb), int(variant)))

    def autotune(self, dtype=np.complex128, nb=1, reps=5):
        """Time the applicable schedules and keep the fastest; returns (variant, ms)."""
        v, ms = C.c_int(), C.c_double()
        check(lib().edk_operator_autotune(self.h, dtype_code(np.dtype(dtype)), int(nb), int(reps),
                                          C.byref(v), C.byref(ms)))
        return v.value, ms.value

    def assemble(self):
        """Assemble (and cache) the sector matrix on the device; returns nnz."""
        n = C.c_int64()
        check(lib().edk_operator_assemble(self.h, C.byref(n)))
        return n.value

    def csr(self):
        """The assembled sector matrix as scipy.sparse.csr_matrix (0-based)."""
        import scipy.sparse as sp
        nnz = self.assemble()
        D = self.B.D
        rp = np.zeros(D + 1, np.int64)
        col = np.zeros(max(nnz, 1), np.int64)
        val = np.zeros(2 * max(nnz, 1))
        check(lib().edk_operator_download_csr(self.h, rp.ctypes.data, col.ctypes.data, val.ctypes.data))
        v = val[: 2 * nnz].view(np.complex128)
        return sp.csr_matrix((v, col[:nnz] - 1, rp - 1), shape=(D, D))

    def release_csr(self):
        check(lib().edk_operator_release_csr(self.h))

    def count_offdiag(self):
        n = C.c_int64()
        check(lib().edk_operator_count_offdiag(self.h, C.byref(n)))
        return n.value

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().edk_operator_destroy(self.h)
            except Exception:
                pass
            self.h = None


def device_operator(H):
    """The cached device handle of H (built on first use)."""
    d = H._dev.get("op")
    if d is None:
        d = DeviceOperator(H)
        H._dev["op"] = d
    return d


def _nb(shape, D):
    if shape[0] != D:
        raise ValueError(f"vector length {shape[0]} does not match the basis dimension {D}")
    return 1 if len(shape) == 1 else int(np.prod(shape[1:]))


def mul_(target, H, v):
    """mul!(target, H, v): target += H v (Operator.jl:309-328)."""
    if is_double(H):
        return double_apply(H, v, target=target, accumulate=True)
    op = device_operator(H)
    D = H.B.D
    if isinstance(v, np.ndarray):
        if not isinstance(target, np.ndarray):
            raise TypeError("target and v must both be host or both be device arrays")
        dt = np.complex64 if v.dtype == np.complex64 and target.dtype == np.complex64 else np.complex128
        x = np.ascontiguousarray(v, dt)
        y = np.ascontiguousarray(target, dt)
        nb = _nb(x.shape, D)
        check(lib().edk_matvec_host(op.h, dtype_code(dt), y.ctypes.data, x.ctypes.data, nb, 1))
        if y is not target:
            target[...] = y
        return target
    nb = _nb(v.shape, D) if hasattr(v, "shape") else 1
    check(lib().edk_matvec(op.h, dtype_code(v.dtype), dev_ptr(target), dev_ptr(v), nb, 1))
    return target


def apply(H, v):
    """H * v with a zero-initialised target of the promoted type (Operator.jl:372-394)."""
    if is_double(H):
        return double_apply(H, v)
    op = device_operator(H)
    D = H.B.D
    if isinstance(v, np.ndarray):
        dt = np.complex64 if v.dtype == np.complex64 else np.complex128
        x = np.ascontiguousarray(v, dt)
        y = np.zeros(x.shape, dt)
        check(lib().edk_matvec_host(op.h, dtype_code(dt), y.ctypes.data, x.ctypes.data,
                                    _nb(x.shape, D), 0))
        return y
    out = DeviceArray(v.shape, v.dtype, ctx=H.B.ctx, zero=False)
    check(lib().edk_matvec(op.h, dtype_code(v.dtype), out.ptr, dev_ptr(v), _nb(v.shape, D), 0))
    return out


def matvec_into(y, H, x, accumulate=False):
    """Device y = (accumulate ? y : 0) + H x without allocation."""
    if is_double(H):
        return double_apply(H, x, target=y, accumulate=accumulate)
    op = device_operator(H)
    nb = _nb(x.shape, H.B.D) if hasattr(x, "shape") else 1
    check(lib().edk_matvec(op.h, dtype_code(x.dtype), dev_ptr(y), dev_ptr(x), nb, int(accumulate)))
    return y


def sparse(H):
    """sparse(opt) (Operator.jl:210-216): the sector matrix, assembled on the device
    (on a DoubleBasis: the rectangular matrix from its columns)."""
    if is_double(H):
        import scipy.sparse as sp
        return sp.csr_matrix(double_array(H))
    return device_operator(H).csr()


def Array(H):
    """Array(opt) (Operator.jl:188-194): the dense sector matrix."""
    if is_double(H):
        return double_array(H)
    return sparse(H).toarray()


def addto_(M, H):
    """addto!(M, opt) (Operator.jl:171-177): accumulate the sector matrix into the
    preallocated M and return M.  A numpy array is updated in place.  A scipy
    CSR or CSC matrix takes over the storage of M + sparse(opt), so the object
    the caller holds carries the sum, as the reference's SparseMatrixCSC does.
    The shapes must agree, and M's element type must hold the matrix's."""
    import scipy.sparse as sp
    if tuple(M.shape) != tuple(H.shape):
        raise ValueError(f"addto_: matrix of shape {tuple(M.shape)} for an operator of shape {tuple(H.shape)}")
    if sp.issparse(M):
        if M.format not in ("csr", "csc"):
            raise TypeError("addto_: a sparse target must be CSR or CSC")
        S = sparse(H)
        if not np.can_cast(S.dtype, M.dtype, casting="same_kind"):
            raise TypeError(f"addto_: cannot accumulate {S.dtype} entries into a {M.dtype} matrix")
        T = (M + S).asformat(M.format)
        T.sum_duplicates()
        M.data, M.indices, M.indptr = T.data.astype(M.dtype, copy=False), T.indices, T.indptr
        M.has_sorted_indices = True
        return M
    M += Array(H)
    return M


# sparse! keeps at most 8 assembled operators, least recently used evicted
# (the reference's LRU{UInt, Any}(maxsize=8), Operator.jl:221)
_SPARSE_CACHE = collections.OrderedDict()
_SPARSE_CACHE_MAX = 8


def sparse_(H):
    """sparse!(opt) (Operator.jl:258-262): assemble and keep the matrix on the device;
    later products with several right-hand sides use it.  Returns the matrix."""
    op = device_operator(H)
    S = op.csr()
    _SPARSE_CACHE[id(op)] = op
    _SPARSE_CACHE.move_to_end(id(op))
    while len(_SPARSE_CACHE) > _SPARSE_CACHE_MAX:
        _, old = _SPARSE_CACHE.popitem(last=False)
        old.release_csr()
    return S


def clear_sparse_cache_(H=None):
    """clear_sparse_cache!() (Operator.jl:272): drop the cached matrices (of H only,
    when given)."""
    if H is not None:
        op = device_operator(H)
        _SPARSE_CACHE.pop(id(op), None)
        op.release_csr()
        return
    while _SPARSE_CACHE:
        _, op = _SPARSE_CACHE.popitem()
        op.release_csr()


def mul(H, v):
    """EDKit.mul(H, v) (Operator.jl:340-370): same result as H * v."""
    return apply(H, v)


def mul_multi(Hs, xs):
    """[H_c @ x_c for c] for up to 8 operators on one basis that share their
    two-site family structure (e.g. XXZ chains of different Delta): one
    launch, every target canonicalised and looked up once for all columns
    (edk_matvec_multi).  xs: DeviceArrays or host vectors of one dtype;
    returns DeviceArrays, or host arrays for host inputs."""
    if not 1 <= len(Hs) == len(xs) <= 8:
        raise ValueError("mul_multi: 1..8 operators, one vector each")
    B = Hs[0].B
    host = not isinstance(xs[0], DeviceArray)
    dx = [DeviceArray.from_host(np.ascontiguousarray(x), ctx=B.ctx) if host else x for x in xs]
    dt = dx[0].dtype
    if any(x.dtype != dt or x.shape != (B.D,) for x in dx):
        raise ValueError("mul_multi: vectors of one dtype and the basis dimension")
    ops = [device_operator(H) for H in Hs]
    dy = [DeviceArray((B.D,), dt, ctx=B.ctx, zero=False) for _ in Hs]
    n = len(Hs)
    P = C.c_void_p * n
    check(lib().edk_matvec_multi(P(*[o.h.value for o in ops]), n, dtype_code(dt), P(*[dev_ptr(y) for y in dy]),
                                 P(*[dev_ptr(x) for x in dx])))
    return [y.to_host() for y in dy] if host else dy

