"""Open-system evolution on the device: lindblad, densitymatrix,
LiouvillianMap, LindbladArnoldiCache, lindblad_timeevolve and
lindblad_timeevolve_ (lindblad_timeevolve!), mirroring
LindbladEvolution.jl:62-611 and Lindblad.jl:11-54 (same keywords, defaults,
diagnostics and error paths).

Density matrices are dense d x d complex128 matrices.  Host (numpy) inputs
are copied to the device column-major (Julia's vec(rho)); the Liouvillian
products run as rocBLAS zgemm on the device and the Arnoldi basis stays in
HBM.  Outputs come back as DensityMatrix objects holding numpy arrays.
"""
import ctypes as C

import numpy as np

from ._lib import LindbladDiag, LindbladParams, check, lib
from .device import DeviceArray, default_context


class DensityMatrix:
    """DensityMatrix (Lindblad.jl:35-38): the matrix rho."""

    def __init__(self, rho):
        self.rho = np.asarray(rho)

    def __array__(self, dtype=None):
        return self.rho if dtype is None else self.rho.astype(dtype)


def densitymatrix(x, L=None, base=2):
    """densitymatrix(rho), densitymatrix(psi) = psi psi^H and
    densitymatrix(i, L; base) (Lindblad.jl:47-54; i is 1-based)."""
    if L is not None:
        n = base ** L
        rho = np.zeros((n, n))
        rho[int(x) - 1, int(x) - 1] = 1
        return DensityMatrix(rho)
    if isinstance(x, DensityMatrix):
        return x
    a = np.asarray(x)
    if a.ndim == 1:
        return DensityMatrix(np.outer(a, a.conj()))
    return DensityMatrix(a)


class Lindblad:
    """lindblad(H, L) (Lindblad.jl:11-33): the Hamiltonian and the jump
    operators, as dense matrices."""

    def __init__(self, H, L):
        self.H = H if hasattr(H, "tocsr") else np.asarray(H)
        self.L = [x if hasattr(x, "tocsr") else np.asarray(x) for x in L]


def lindblad(H, L):
    return Lindblad(H, L)


def _dense(a):
    """Dense form of a matrix argument (scipy sparse matrices are densified)."""
    return a.toarray() if hasattr(a, "toarray") else np.asarray(a)


def _colmajor(a):
    """Device layout of a matrix: column-major, complex128."""
    return np.ascontiguousarray(np.asarray(_dense(a), np.complex128).T)


def _csr(a):
    """0-based CSR arrays (int64 row pointers, int32 columns, interleaved complex values)."""
    import scipy.sparse as sp
    m = sp.csr_matrix(a, dtype=np.complex128)
    m.sort_indices()
    vals = np.ascontiguousarray(m.data.astype(np.complex128)).view(np.float64)
    return (np.ascontiguousarray(m.indptr.astype(np.int64)), np.ascontiguousarray(m.indices.astype(np.int32)),
            np.ascontiguousarray(vals))


def _rho_of(dm):
    return np.asarray(dm.rho if isinstance(dm, DensityMatrix) else dm)


class LiouvillianMap:
    """LiouvillianMap(H, jumps) or LiouvillianMap(lb) (LindbladEvolution.jl:62-92)."""

    def __init__(self, H, jumps=None, ctx=None, sparse=None):
        """sparse: None follows the input (scipy sparse H selects the sparse
        form, as issparse(H) does in the reference); True/False forces it."""
        if isinstance(H, Lindblad):
            H, jumps = H.H, H.L
        jumps = [] if jumps is None else list(jumps)
        if sparse is None:
            sparse = hasattr(H, "tocsr")
        if not hasattr(H, "tocsr"):
            H = np.asarray(H)
        jumps = [L if hasattr(L, "tocsr") else np.asarray(L) for L in jumps]
        shape = H.shape
        if len(shape) != 2 or shape[0] != shape[1]:
            raise ValueError("LiouvillianMap: H must be square")
        d = shape[0]
        for j, L in enumerate(jumps):
            if tuple(L.shape) != (d, d):
                raise ValueError(f"LiouvillianMap: jump {j + 1} has incompatible size")
        self.ctx = ctx or default_context()
        self.dim = d
        self.njumps = len(jumps)
        self.sparse = bool(sparse)
        h = C.c_void_p()
        if self.sparse:
            hp, hc, hv = _csr(H)
            keep = [hp, hc, hv]
            if jumps:
                parts = [_csr(L) for L in jumps]
                offs = np.cumsum([0] + [p[1].size for p in parts])
                jp = np.ascontiguousarray(np.concatenate([p[0] + o for p, o in zip(parts, offs[:-1])]))
                jc = np.ascontiguousarray(np.concatenate([p[1] for p in parts]))
                jv = np.ascontiguousarray(np.concatenate([p[2] for p in parts]))
                keep += [jp, jc, jv]
                jargs = (jp.ctypes.data, jc.ctypes.data, jv.ctypes.data)
            else:
                jargs = (None, None, None)
            check(lib().edk_liouvillian_create_csr(self.ctx.h, d, hp.ctypes.data, hc.ctypes.data, hv.ctypes.data,
                                                   len(jumps), *jargs, C.byref(h)))
            del keep
        else:
            dH = DeviceArray.from_host(_colmajor(H), ctx=self.ctx)
            dJ = (DeviceArray.from_host(np.stack([_colmajor(L) for L in jumps]), ctx=self.ctx) if jumps else None)
            check(lib().edk_liouvillian_create(self.ctx.h, d, dH.ptr, len(jumps), dJ.ptr if dJ else None,
                                               C.byref(h)))
        self.h = h

    @property
    def shape(self):
        return (self.dim ** 2, self.dim ** 2)

    def apply_device(self, x, y):
        """y = L(x) for device d x d buffers (column-major)."""
        check(lib().edk_liouvillian_apply(self.h, x.ptr, y.ptr))
        return y

    def __matmul__(self, rho):
        """A * rho for a d x d matrix or A * v for vec(rho) (LindbladEvolution.jl:138-151)."""
        a = np.asarray(rho, np.complex128)
        d = self.dim
        vecin = a.ndim == 1
        if vecin:
            if a.size != d * d:
                raise ValueError("LiouvillianMap: vector has incompatible length")
            a = a.reshape(d, d, order="F")
        elif a.shape != (d, d):
            raise ValueError("LiouvillianMap: matrix has incompatible size")
        x = DeviceArray.from_host(_colmajor(a), ctx=self.ctx)
        y = DeviceArray((d, d), np.complex128, ctx=self.ctx, zero=False)
        self.apply_device(x, y)
        out = y.to_host().T
        return out.reshape(-1, order="F") if vecin else out

    __mul__ = __matmul__

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().edk_liouvillian_destroy(self.h)
            except Exception:
                pass
            self.h = None


def _as_map(A):
    return A if isinstance(A, LiouvillianMap) else LiouvillianMap(A)


class LindbladArnoldiCache:
    """LindbladArnoldiCache(A, dm; tol, m_init, m_max, extend_step, nsample,
    bisect_iters, extend_basis, normalize_trace, record_min_eig)
    (LindbladEvolution.jl:154-273)."""

    def __init__(self, A, dm, tol=1e-10, m_init=20, m_max=50, extend_step=10, nsample=17, bisect_iters=12,
                 extend_basis=True, normalize_trace=False, record_min_eig=False):
        self.A = _as_map(A)
        d = self.A.dim
        rho = _rho_of(densitymatrix(dm))
        if rho.shape != (d, d):
            raise ValueError("LindbladArnoldiCache: density matrix has incompatible size")
        self.dim = d
        self.rho0 = DeviceArray.from_host(_colmajor(rho), ctx=self.A.ctx)
        p = LindbladParams(float(tol), int(m_init), int(m_max), int(extend_step), int(nsample), int(bisect_iters),
                           int(bool(extend_basis)), int(bool(normalize_trace)), int(bool(record_min_eig)))
        h = C.c_void_p()
        check(lib().edk_lindblad_create(self.A.h, self.rho0.ptr, C.byref(p), C.byref(h)))
        self.h = h

    def evolve_device(self, t, out=None):
        """lindblad_timeevolve!(cache, t) into a device d x d buffer (column-major)."""
        out = out if out is not None else DeviceArray((self.dim, self.dim), np.complex128, ctx=self.A.ctx,
                                                      zero=False)
        check(lib().edk_lindblad_evolve(self.h, float(t), out.ptr))
        return out

    def evolve(self, t):
        return DensityMatrix(self.evolve_device(t).to_host().T)

    @property
    def m(self):
        m = C.c_int()
        check(lib().edk_lindblad_dim(self.h, C.byref(m)))
        return m.value

    @property
    def diagnostics(self):
        """LindbladArnoldiDiagnostics (LindbladEvolution.jl:17-37) as a dict."""
        d = LindbladDiag()
        check(lib().edk_lindblad_diagnostics(self.h, C.byref(d)))
        out = {k: getattr(d, k) for k, _ in LindbladDiag._fields_}
        na, no, ne = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().edk_lindblad_records(self.h, None, None, None, None, 0, C.byref(na), C.byref(no),
                                         C.byref(ne)))
        cap = max(1, na.value, no.value, ne.value)
        bufs = [np.zeros(cap) for _ in range(4)]
        check(lib().edk_lindblad_records(self.h, *[b.ctypes.data for b in bufs], cap, C.byref(na), C.byref(no),
                                         C.byref(ne)))
        out["accepted_intervals"] = bufs[0][: na.value].tolist()
        out["trace_drifts"] = bufs[1][: no.value].tolist()
        out["hermiticity_drifts"] = bufs[2][: no.value].tolist()
        out["min_hermitian_eigs"] = bufs[3][: ne.value].tolist()
        return out

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().edk_lindblad_destroy(self.h)
            except Exception:
                pass
            self.h = None


def lindblad_timeevolve_(cache, t):
    """lindblad_timeevolve!(cache, t) (LindbladEvolution.jl:598-611)."""
    return cache.evolve(t)


def lindblad_timeevolve(A, dm, t, return_diagnostics=False, **kw):
    """lindblad_timeevolve(A, dm, t) or (A, dm, ts) (LindbladEvolution.jl:537-596);
    A: a Lindblad or a LiouvillianMap.  Times are served in sorted order and
    returned in the given order."""
    scalar = np.isscalar(t)
    ts = np.atleast_1d(np.asarray(t, np.float64))
    if np.any(ts < 0):
        if scalar:
            raise ValueError("lindblad_timeevolve: t must be non-negative; backward evolution is not supported")
        raise ValueError("lindblad_timeevolve: all times in ts must be non-negative; backward evolution is not supported")
    cache = LindbladArnoldiCache(A, dm, **kw)
    out = [None] * len(ts)
    for i in np.argsort(ts, kind="stable"):
        out[i] = cache.evolve(ts[i])
    res = out[0] if scalar else out
    return (res, cache.diagnostics) if return_diagnostics else res
