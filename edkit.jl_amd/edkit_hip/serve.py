"""Serving loop: Lanczos ground-state and Krylov time-evolution jobs on one
device, admitted by worker slot and HBM footprint (libedkit_hip's
edk_server_*).  Jobs of any sector and Hamiltonian run side by side on their
own streams; jobs sharing a sector and term structure share their matvecs,
one block launch per Krylov step; each result is the same alone and in a mix.
"""
import ctypes as C

import numpy as np

from ._lib import (EDK_JOB_EVOLVE, EDK_JOB_GROUND, REORTH_CODES, Job, JobResult, ServerInfo, ServerParams,
                   check, lib)
from .device import dtype_code
from .matvec import device_operator


class Server:
    """edk_server handle: `workers` threads (one stream each), admission
    budget `hbm_budget` bytes (0: 80 % of free device memory).  batch: the
    matvecs of jobs sharing a sector and term structure run as one block
    launch per Krylov step (up to 8 jobs); batch_wait_us: longest wait for a
    group's other members; affinity: admit queued jobs that join a running
    group first."""

    def __init__(self, workers=8, hbm_budget=0, device=0, batch=True, batch_wait_us=250.0, affinity=True):
        p = ServerParams()
        lib().edk_server_default_params(C.byref(p))
        p.batch, p.batch_wait_us, p.affinity = int(batch), float(batch_wait_us), int(affinity)
        h = C.c_void_p()
        check(lib().edk_server_create_ex(int(device), int(workers), int(hbm_budget), C.byref(p), C.byref(h)))
        self.h = h
        self._pending = {}

    def _submit(self, H, kind, dtype, psi0, psi0_index, **kw):
        op = device_operator(H)
        dt = np.dtype(dtype)
        keep = None
        ptr = None
        if psi0 is not None:
            keep = np.ascontiguousarray(psi0, dt)
            if keep.shape != (H.B.D,):
                raise ValueError("start vector length does not match the basis dimension")
            ptr = keep.ctypes.data
        job = Job(kind, dtype_code(dt), op.h.value, ptr, int(psi0_index), int(kw.get("m", 0)),
                  int(kw.get("refine", 0)), float(kw.get("tol", 0.0)), float(kw.get("t", 0.0)),
                  int(kw.get("m_init", 0)), int(kw.get("m_max", 0)), int(kw.get("want_vector", 0)),
                  REORTH_CODES[kw.get("reorth", "full")])
        tk = C.c_int64()
        check(lib().edk_server_submit(self.h, C.byref(job), C.byref(tk)))
        # the operator must outlive the job; the start vector was copied
        self._pending[tk.value] = (H, op, dt, H.B.D, bool(kw.get("want_vector", 0)))
        return tk.value

    def submit_ground(self, H, psi0=None, psi0_index=0, m=300, tol=1e-9, refine=False,
                      dtype=np.complex128, want_vector=False, reorth="full"):
        """Lanczos ground state of H from psi0 (host vector) or the product
        state at psi0_index; returns a ticket.  reorth: "full" or "partial"
        (as in lanczos_ground)."""
        if reorth not in REORTH_CODES:
            raise ValueError("reorth must be 'full' or 'partial'")
        return self._submit(H, EDK_JOB_GROUND, dtype, psi0, psi0_index, m=m, tol=tol,
                            refine=int(refine), want_vector=int(want_vector), reorth=reorth)

    def submit_evolve(self, H, t, psi0=None, psi0_index=0, tol=1e-10, m_init=30, m_max=60,
                      dtype=np.complex128, want_vector=False, reorth="full"):
        """exp(-iHt) psi0 through KrylovEvolutionCache semantics; returns a
        ticket.  reorth "partial" keeps the Krylov basis semi-orthogonal
        (Simon's omega recurrence) instead of the reference's full sweeps."""
        if reorth not in REORTH_CODES:
            raise ValueError("reorth must be 'full' or 'partial'")
        return self._submit(H, EDK_JOB_EVOLVE, dtype, psi0, psi0_index, t=t, tol=tol,
                            m_init=m_init, m_max=m_max, want_vector=int(want_vector), reorth=reorth)

    def wait(self, ticket, timeout=-1.0):
        """Block until the job is done; returns (result dict, vector or None)."""
        H, op, dt, D, want = self._pending[ticket]
        res = JobResult()
        vec = np.empty(D, dt) if want else None
        check(lib().edk_server_wait(self.h, int(ticket), float(timeout), C.byref(res),
                                    vec.ctypes.data if want else None))
        del self._pending[ticket]
        out = {f: getattr(res, f) for f, _ in JobResult._fields_ if f != "message"}
        out["message"] = res.message.decode(errors="replace")
        return out, vec

    def clock(self):
        return lib().edk_server_clock(self.h)

    def stats(self):
        st = ServerInfo()
        check(lib().edk_server_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerInfo._fields_}

    def batch_stats(self, weighted=False):
        """(block launches, columns carried) of the batched matvecs so far;
        weighted=True adds the columns per launch weighted by the sector
        dimension."""
        a, b, w = C.c_int64(), C.c_int64(), C.c_double()
        check(lib().edk_server_batch_stats(self.h, C.byref(a), C.byref(b), C.byref(w)))
        return (a.value, b.value, w.value) if weighted else (a.value, b.value)

    def close(self):
        if getattr(self, "h", None):
            lib().edk_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
