"""ctypes wrapper of the CPU oracle (oracle/edk_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as the product path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

KIND = {"tensor": 0, "projected": 1, "translational": 2,
        "translation_parity": 3, "abelian": 4, "parity": 5, "flip": 6, "parity_flip": 7,
        "translation_flip": 8}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libedk_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        P = C.c_void_p
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
        u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
        L.orc_basis_build.restype = P
        L.orc_basis_build.argtypes = [C.c_int] * 11 + [C.c_int, P, P, P, C.c_int]
        L.orc_basis_free.argtypes = [P]
        L.orc_basis_dim.restype = C.c_int64
        L.orc_basis_dim.argtypes = [P]
        L.orc_basis_order.restype = C.c_int
        L.orc_basis_order.argtypes = [P]
        L.orc_basis_tables.argtypes = [P, i64p, f64p]
        L.orc_basis_index.argtypes = [P, u64p, C.c_int64, i64p, f64p, f64p]
        L.orc_schmidt.argtypes = [P, f64p, C.c_int, i32p, f64p]
        L.orc_op_create.restype = P
        L.orc_op_create.argtypes = [P, C.c_int64, i64p, i64p, i64p, i64p, i64p, i64p, f64p]
        L.orc_op_free.argtypes = [P]
        L.orc_mul.argtypes = [P, f64p, f64p, C.c_int]
        L.orc_mul_block.argtypes = [P, f64p, f64p, C.c_int64, C.c_int]
        L.orc_mul_cols.argtypes = [P, f64p, f64p, i64p, C.c_int64, C.c_int]
        L.orc_column.restype = C.c_int64
        L.orc_column.argtypes = [P, C.c_int64, i64p, f64p, C.c_int64]
        L.orc_count_offdiag.restype = C.c_int64
        L.orc_count_offdiag.argtypes = [P, C.c_int]
        L.orc_tridiag_eig.restype = C.c_int
        L.orc_tridiag_eig.argtypes = [C.c_int, f64p, f64p, f64p, f64p]
        L.orc_lanczos_ground.restype = C.c_int
        L.orc_lanczos_ground.argtypes = [P, f64p, C.c_int, C.c_double, C.c_int,
                                         C.POINTER(C.c_double), C.POINTER(C.c_double), P]
        L.orc_krylov_create.restype = P
        L.orc_krylov_create.argtypes = [P, f64p, C.c_double] + [C.c_int] * 8
        L.orc_krylov_evolve.restype = C.c_int
        L.orc_krylov_evolve.argtypes = [P, C.c_double, f64p]
        L.orc_krylov_diagnostics.argtypes = [P, i64p]
        L.orc_krylov_m.restype = C.c_int
        L.orc_krylov_m.argtypes = [P]
        L.orc_krylov_free.argtypes = [P]
        L.orc_last_error.restype = C.c_char_p
        L.orc_gosper_next.restype = C.c_uint64
        L.orc_gosper_next.argtypes = [C.c_uint64]
        L.orc_benes_apply.restype = C.c_uint64
        L.orc_benes_apply.argtypes = [i32p, C.c_int, C.c_uint64]
        L.orc_perm_apply_int.restype = C.c_uint64
        L.orc_perm_apply_int.argtypes = [i32p, C.c_int, C.c_uint64]
        L.orc_compute_period.restype = C.c_int
        L.orc_compute_period.argtypes = [i32p, u8p, C.c_int]
        _LIB = L
    return _LIB


def _err():
    return lib().orc_last_error().decode()


def _cview(z):
    z = np.ascontiguousarray(z, dtype=np.complex128)
    return z, z.view(np.float64)


class Basis:
    """Oracle basis.  kind in KIND; keywords as the reference constructors."""

    def __init__(self, kind, L, base=2, a=1, N=None, k=0, p=1, z=0,
                 has_k=False, has_p=False, has_z=False, symmetries=None,
                 nthreads=None):
        nthreads = nthreads or min(16, os.cpu_count() or 1)
        perms = ks = invs = None
        nc = 0
        keep = []
        if symmetries:
            nc = len(symmetries)
            perms = np.ascontiguousarray(
                np.concatenate([np.asarray(s[0], np.int32) for s in symmetries]))
            ks = np.ascontiguousarray(np.asarray([s[1] for s in symmetries], np.int32))
            invs = np.ascontiguousarray(np.concatenate(
                [np.asarray(s[2] if len(s) > 2 else [0] * L, np.uint8) for s in symmetries]))
            keep = [perms, ks, invs]
        ptr = lambda a: a.ctypes.data if a is not None else None  # noqa: E731
        self.h = lib().orc_basis_build(
            KIND[kind], L, base, a, -1 if N is None else N, k, p, z,
            int(has_k), int(has_p), int(has_z), nc, ptr(perms), ptr(ks), ptr(invs),
            nthreads)
        del keep
        if not self.h:
            raise ValueError(_err())
        self.kind, self.L, self.base = kind, L, base
        self.D = lib().orc_basis_dim(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_basis_free(self.h)
            self.h = None

    @property
    def order(self):
        return lib().orc_basis_order(self.h)

    def tables(self):
        I = np.zeros(self.D, np.int64)
        R = np.zeros(self.D, np.float64)
        lib().orc_basis_tables(self.h, I, R)
        return I, R

    def index(self, states):
        states = np.ascontiguousarray(states, dtype=np.uint64)
        n = states.size
        pos = np.zeros(n, np.int64)
        cre = np.zeros(n)
        cim = np.zeros(n)
        lib().orc_basis_index(self.h, states, n, pos, cre, cim)
        return pos, cre + 1j * cim

    def schmidt(self, v, Ainds):
        """schmidt(v, Ainds, b) (Schmidt.jl:186-330): the dA x dB matrix."""
        x = np.ascontiguousarray(np.asarray(v, np.complex128))
        A = np.ascontiguousarray(np.asarray(list(Ainds), np.int32))
        dA, dB = self.base ** len(A), self.base ** (self.L - len(A))
        M = np.zeros(dA * dB, np.complex128)
        lib().orc_schmidt(self.h, x.view(np.float64), len(A), A, M.view(np.float64))
        return M.reshape((dB, dA)).T


class Op:
    """Oracle operator over a flattened term list (see Operator.flatten())."""

    def __init__(self, basis, flat):
        self.basis = basis
        self._flat = [np.ascontiguousarray(a) for a in flat[1:]]
        so, si, co, cp, no, rv, nz = self._flat
        self.h = lib().orc_op_create(basis.h, flat[0], so, si, co, cp, no, rv,
                                     nz.astype(np.float64).reshape(-1))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_op_free(self.h)
            self.h = None

    def mul_(self, target, v, nthreads=1):
        """mul!(target, H, v): accumulates into target (complex128)."""
        t, tv = _cview(target)
        _, vv = _cview(v)
        if v.ndim == 2:
            lib().orc_mul_block(self.h, tv.reshape(-1), vv.reshape(-1), v.shape[1], nthreads)
        else:
            lib().orc_mul(self.h, tv, vv, nthreads)
        if t is not target:
            target[...] = t
        return target

    def __matmul__(self, v):
        out = np.zeros(v.shape, np.complex128)
        return self.mul_(out, v)

    def mul_cols_(self, target, v, cols, nthreads):
        """Threaded mul restricted to the given columns, accumulated into target."""
        t, tv = _cview(target)
        _, vv = _cview(v)
        lib().orc_mul_cols(self.h, tv, vv, np.ascontiguousarray(cols, np.int64), len(cols), nthreads)
        if t is not target:
            target[...] = t
        return target

    def column(self, j, cap=4096):
        """Column j (0-based) of the sector matrix: (0-based rows, values)."""
        pos = np.zeros(cap, np.int64)
        val = np.zeros(2 * cap)
        n = lib().orc_column(self.h, int(j), pos, val, cap)
        if n < 0:
            raise ValueError("column has more entries than cap")
        return pos[:n] - 1, val[0:2 * n:2] + 1j * val[1:2 * n:2]

    def count_offdiag(self, nthreads=None):
        return lib().orc_count_offdiag(self.h, nthreads or min(16, os.cpu_count() or 1))


def tridiag_eig(alpha, beta):
    m = len(alpha)
    lam = np.zeros(m)
    Q = np.zeros(m * m)
    rc = lib().orc_tridiag_eig(m, np.ascontiguousarray(alpha, np.float64),
                               np.ascontiguousarray(np.append(beta, 0.0), np.float64), lam, Q)
    if rc:
        raise RuntimeError(_err())
    return lam, Q.reshape(m, m).T.copy()


def lanczos_ground(op, v0, m, tol=0.0, nthreads=1, want_vec=False):
    _, vv = _cview(v0)
    E0 = C.c_double()
    res = C.c_double()
    vec = np.zeros(op.basis.D, np.complex128) if want_vec else None
    steps = lib().orc_lanczos_ground(op.h, vv, m, tol, nthreads, C.byref(E0), C.byref(res),
                                     vec.ctypes.data if want_vec else None)
    if steps < 0:
        raise RuntimeError(_err())
    return E0.value, res.value, steps, vec


class Krylov:
    """KrylovEvolutionCache restatement."""

    def __init__(self, op, psi0, tol=1e-10, m_init=30, m_max=60, extend_step=10,
                 nsample=13, bisect_iters=12, extend_basis=True,
                 normalize_output=False, nthreads=1):
        self.op = op
        _, pv = _cview(psi0)
        self.h = lib().orc_krylov_create(op.h, pv, tol, m_init, m_max, extend_step,
                                         nsample, bisect_iters, int(extend_basis),
                                         int(normalize_output), nthreads)
        if not self.h:
            raise ValueError(_err())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_krylov_free(self.h)
            self.h = None

    def evolve(self, t):
        out = np.zeros(self.op.basis.D, np.complex128)
        if lib().orc_krylov_evolve(self.h, float(t), out.view(np.float64)):
            raise ValueError(_err())
        return out

    @property
    def m(self):
        return lib().orc_krylov_m(self.h)

    def diagnostics(self):
        d = np.zeros(7, np.int64)
        lib().orc_krylov_diagnostics(self.h, d)
        keys = ["basis_builds", "basis_extensions", "restarts", "matvecs",
                "total_times_served", "max_dim_used", "n_accepted"]
        return dict(zip(keys, d.tolist()))


def timeevolve(op, psi0, ts, **kw):
    """timeevolve(H, psi0, ts) (TimeEvolution.jl:563-577): sorts ts."""
    ts = np.atleast_1d(np.asarray(ts, np.float64))
    if np.any(ts < 0):
        raise ValueError("timeevolve: all times in ts must be non-negative")
    c = Krylov(op, psi0, **kw)
    perm = np.argsort(ts, kind="stable")
    out = np.zeros((op.basis.D, len(ts)), np.complex128)
    for i in perm:
        out[:, i] = c.evolve(ts[i])
    return out, c.diagnostics()


def gosper_next(s):
    return lib().orc_gosper_next(s)


def benes_apply(perm, L, s):
    return lib().orc_benes_apply(np.ascontiguousarray(perm, np.int32), L, s)


def perm_apply_int(perm, L, s):
    return lib().orc_perm_apply_int(np.ascontiguousarray(perm, np.int32), L, s)


def compute_period(perm, inv=None):
    L = len(perm)
    inv = np.zeros(L, np.uint8) if inv is None else np.ascontiguousarray(inv, np.uint8)
    return lib().orc_compute_period(np.ascontiguousarray(perm, np.int32), inv, L)
