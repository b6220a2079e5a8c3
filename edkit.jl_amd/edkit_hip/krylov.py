"""Device Lanczos ground state and Krylov time evolution.

KrylovEvolutionCache / timeevolve / timeevolve! mirror TimeEvolution.jl:84-649
(same keywords, defaults, diagnostics and error paths); lanczos_ground is a
new entry point built from the same _lanczos_build! recurrence.
"""
import ctypes as C

import numpy as np

from ._lib import REORTH_CODES, KrylovDiag, KrylovParams, check, lib
from .device import DeviceArray, dev_ptr, dtype_code
from .matvec import device_operator


def _to_device(v, B, dtype):
    if isinstance(v, np.ndarray):
        return DeviceArray.from_host(np.asarray(v, dtype), ctx=B.ctx)
    return v


def lanczos_ground(H, v0, m=300, tol=0.0, dtype=np.complex128, refine=True, want_vec=False,
                   reorth="full", info=None):
    """Lowest eigenpair by Lanczos.  Returns (E0, residual, steps, vec or None).
    dtype complex64 stores the Krylov vectors and runs the matvec in FP32
    with FP64 scalars; refine adds the final FP64 Rayleigh quotient/residual
    on the Ritz vector.  reorth: "full" reorthogonalises every step against
    the whole basis; "partial" only when the omega-recurrence estimate of the
    loss of orthogonality passes sqrt(eps) (Simon 1984).  info (a dict, if
    given) receives "reorth_steps"."""
    if reorth not in REORTH_CODES:
        raise ValueError("reorth must be 'full' or 'partial'")
    op = device_operator(H)
    B = H.B
    x0 = _to_device(v0, B, dtype)
    dt = x0.dtype
    vec = DeviceArray((B.D,), dt, ctx=B.ctx, zero=False) if want_vec else None
    E0, res, steps, nre = C.c_double(), C.c_double(), C.c_int(), C.c_int()
    check(lib().edk_lanczos_ground_ex(op.h, dtype_code(dt), dev_ptr(x0), int(m), float(tol),
                                      int(refine), REORTH_CODES[reorth], C.byref(E0), C.byref(res),
                                      C.byref(steps), C.byref(nre),
                                      vec.ptr if vec is not None else None))
    if info is not None:
        info["reorth_steps"] = nre.value
    return E0.value, res.value, steps.value, vec


class KrylovEvolutionCache:
    """KrylovEvolutionCache(H, psi0; tol, m_init, m_max, extend_step, nsample,
    bisect_iters, reuse_basis, extend_basis, normalize_output)."""

    def __init__(self, H, psi0, tol=1e-10, m_init=30, m_max=60, extend_step=10, nsample=13,
                 bisect_iters=12, reuse_basis=True, extend_basis=True, normalize_output=False,
                 dtype=np.complex128, reorth="full"):
        self.H = H
        self.op = device_operator(H)
        self.B = H.B
        self.psi0 = _to_device(psi0, self.B, dtype)
        self.dtype = self.psi0.dtype
        if reorth not in REORTH_CODES:
            raise ValueError("reorth must be 'full' or 'partial'")
        p = KrylovParams(float(tol), int(m_init), int(m_max), int(extend_step), int(nsample),
                         int(bisect_iters), int(bool(extend_basis)), int(bool(normalize_output)),
                         REORTH_CODES[reorth])
        h = C.c_void_p()
        check(lib().edk_krylov_create(self.op.h, dtype_code(self.dtype), dev_ptr(self.psi0),
                                      C.byref(p), C.byref(h)))
        self.h = h

    def evolve(self, t, out=None):
        """timeevolve!(cache, t) -> device vector."""
        out = out if out is not None else DeviceArray((self.B.D,), self.dtype, ctx=self.B.ctx, zero=False)
        check(lib().edk_krylov_evolve(self.h, float(t), dev_ptr(out)))
        return out

    @property
    def diagnostics(self):
        d = KrylovDiag()
        check(lib().edk_krylov_diagnostics(self.h, C.byref(d)))
        out = {k: getattr(d, k) for k, _ in KrylovDiag._fields_}
        n = C.c_int64()
        buf = np.zeros(max(1, out["n_accepted_intervals"]))
        check(lib().edk_krylov_accepted_intervals(self.h, buf.ctypes.data, buf.size, C.byref(n)))
        out["accepted_intervals"] = buf[: n.value].tolist()
        return out

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().edk_krylov_destroy(self.h)
            except Exception:
                pass
            self.h = None


def timeevolve_(cache, t):
    """timeevolve!(cache, t) (TimeEvolution.jl:586-597), returned on the host."""
    return cache.evolve(t).to_host()


def timeevolve(H, psi0, t, return_diagnostics=False, dtype=np.complex128, **kw):
    """timeevolve(H, psi0, t) or timeevolve(H, psi0, ts) (TimeEvolution.jl:556-577).
    Host input gives host output; ts are served in sorted order."""
    scalar = np.isscalar(t)
    ts = np.atleast_1d(np.asarray(t, np.float64))
    if not scalar and np.any(ts < 0):
        raise ValueError("timeevolve: all times in ts must be non-negative; backward evolution is not supported")
    cache = KrylovEvolutionCache(H, psi0, dtype=dtype, **kw)
    perm = np.argsort(ts, kind="stable")
    out = np.zeros((H.B.D, len(ts)), cache.dtype)
    for i in perm:
        out[:, i] = cache.evolve(ts[i]).to_host()
    res = out[:, 0] if scalar else out
    return (res, cache.diagnostics) if return_diagnostics else res
