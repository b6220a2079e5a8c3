"""CPU restatement of the adaptive Arnoldi Lindblad propagator
(LindbladEvolution.jl:62-611) and of the explicit Liouvillian superoperator
the reference's own tests compare against
(test/lindblad_timeevolve_tests.jl:4-26).

TEST INFRASTRUCTURE ONLY.  Imported by tests/ and by bench.py's cpu_baseline
leg, always as the checker or the timed CPU baseline, never as the product
path.  Plain numpy/scipy, column-major vec(rho) throughout (Julia's vec).
"""
import numpy as np
import scipy.linalg as sla


def vec(rho):
    """Julia vec: column-major flattening."""
    return np.asarray(rho).reshape(-1, order="F")


def unvec(v, d):
    return np.asarray(v).reshape(d, d, order="F")


def explicit_liouvillian(H, jumps):
    """The d^2 x d^2 superoperator from vec(AXB) = (B^T kron A) vec(X)
    (lindblad_timeevolve_tests.jl:4-19)."""
    H = np.asarray(H, np.complex128)
    d = H.shape[0]
    I = np.eye(d, dtype=np.complex128)
    D = np.zeros((d, d), np.complex128)
    for L in jumps:
        D += L.conj().T @ L
    M = -1j * (np.kron(I, H) - np.kron(H.T, I))
    for L in jumps:
        M += np.kron(L.conj(), L)
    M += -0.5 * np.kron(I, D)
    M += -0.5 * np.kron(D.T, I)
    return M


def explicit_evolve(H, jumps, rho0, t):
    """exp(t L) vec(rho0) (lindblad_timeevolve_tests.jl:21-25)."""
    d = rho0.shape[0]
    return unvec(sla.expm(t * explicit_liouvillian(H, jumps)) @ vec(rho0), d)


def liouvillian_apply(H, jumps, X):
    """_liouvillian_apply! (LindbladEvolution.jl:114-136), term by term."""
    D = sum((L.conj().T @ L for L in jumps), np.zeros_like(H))
    Y = -1j * (H @ X - X @ H)
    for L in jumps:
        Y = Y + (L @ X) @ L.conj().T
    Y = Y + -0.5 * (D @ X)
    Y = Y + -0.5 * (X @ D)
    return Y


class ArnoldiCache:
    """LindbladArnoldiCache (LindbladEvolution.jl:176-273) with its
    controller (:282-535)."""

    def __init__(self, H, jumps, rho0, tol=1e-10, m_init=20, m_max=50, extend_step=10, nsample=17,
                 bisect_iters=12, extend_basis=True, normalize_trace=False, record_min_eig=False):
        self.H = np.asarray(H, np.complex128)
        self.jumps = [np.asarray(L, np.complex128) for L in jumps]
        d = self.H.shape[0]
        rho0 = np.asarray(rho0, np.complex128)
        if rho0.shape != (d, d):
            raise ValueError("LindbladArnoldiCache: density matrix has incompatible size")
        if m_init < 1:
            raise ValueError("m_init must be >= 1")
        if m_max < m_init:
            raise ValueError("m_max must be >= m_init")
        if extend_step < 1:
            raise ValueError("extend_step must be >= 1")
        if nsample < 2:
            raise ValueError("nsample must be >= 2")
        if bisect_iters < 0:
            raise ValueError("bisect_iters must be >= 0")
        N = d * d
        self.d, self.N = d, N
        self.m_max = min(int(m_max), N)
        self.m_init = min(int(m_init), self.m_max)
        self.tol, self.extend_step, self.nsample = float(tol), int(extend_step), int(nsample)
        self.bisect_iters, self.extend_basis = int(bisect_iters), bool(extend_basis)
        self.normalize_trace, self.record_min_eig = bool(normalize_trace), bool(record_min_eig)
        self.t_anchor = 0.0
        self.t_latest = 0.0
        self.rho_anchor = rho0.copy()
        nrm = np.linalg.norm(self.rho_anchor)
        if not nrm > 0:
            raise ValueError("LindbladArnoldiCache: initial density matrix has zero Frobenius norm")
        self.norm_anchor = nrm
        self.trace_target = np.trace(self.rho_anchor)
        self.V = np.zeros((N, self.m_max + 1), np.complex128)
        self.V[:, 0] = vec(self.rho_anchor) / nrm
        self.Hred = np.zeros((self.m_max + 1, self.m_max), np.complex128)
        self.m = 0
        self.step_cache = {}
        self.diag = dict(basis_builds=0, basis_extensions=0, restarts=0, liouvillian_applies=0,
                         total_times_served=0, max_dim_used=0, accepted_intervals=[], trace_drifts=[],
                         hermiticity_drifts=[], min_hermitian_eigs=[], max_trace_drift=0.0,
                         max_hermiticity_drift=0.0, min_min_hermitian_eig=np.inf)
        self._arnoldi_build(self.m_init)
        self._build_reduced()
        self.diag["basis_builds"] += 1
        self.diag["max_dim_used"] = max(self.diag["max_dim_used"], self.m)

    def _arnoldi_build(self, m_target):
        """_arnoldi_build! (:282-321): two modified Gram-Schmidt sweeps."""
        m_target = min(int(m_target), self.m_max)
        V, Hr = self.V, self.Hred
        for j in range(self.m + 1, m_target + 1):
            w = vec(liouvillian_apply(self.H, self.jumps, unvec(V[:, j - 1], self.d)))
            self.diag["liouvillian_applies"] += 1
            Hr[:, j - 1] = 0
            for _ in range(2):
                for i in range(j):
                    hij = np.vdot(V[:, i], w)
                    Hr[i, j - 1] += hij
                    w = w - hij * V[:, i]
            hnext = np.linalg.norm(w)
            Hr[j, j - 1] = hnext
            self.m = j
            if hnext <= np.sqrt(np.finfo(float).eps) * max(1.0, np.linalg.norm(Hr[:j, j - 1])):
                Hr[j, j - 1] = 0
                return True
            V[:, j] = w / hnext
        return False

    def _build_reduced(self):
        """_build_reduced! (:323-336): complex Schur form of H[1:m, 1:m]."""
        m = self.m
        T, Z = sla.schur(self.Hred[:m, :m], output="complex")
        self.T, self.Z = T, Z
        self.q1 = Z[0, :].conj()
        self.rowm = Z[m - 1, :].copy()
        self.step_cache = {}

    def _boundary(self):
        return 0.0 if self.m < 1 else self.Hred[self.m, self.m - 1]

    def _step_matrix(self, tau):
        key = float(tau)
        if key not in self.step_cache:
            self.step_cache[key] = np.triu(sla.expm(self.T * key))
        return self.step_cache[key]

    def _reduced_y(self, tau):
        if tau == 0:
            return self.q1.copy()
        return self._step_matrix(tau) @ self.q1

    def _defect_from_y(self, y):
        h = self._boundary()
        if h == 0:
            return 0.0
        return float(self.norm_anchor * abs(h) * abs(np.dot(self.rowm, y)))

    def _effective_nsample(self, tau):
        base = self.nsample
        if self.m <= 1:
            return base
        g = np.linalg.norm(self.Hred[:self.m, :self.m], 1)
        if g <= 0:
            return base
        n = int(np.ceil(4 * g * tau)) + 1
        return int(min(max(base, n), 513))

    def _max_valid_interval(self, tau):
        """_max_valid_interval (:387-421)."""
        if not tau > 0:
            return 0.0
        if self._boundary() == 0:
            return float(tau)
        n = self._effective_nsample(tau)
        xs = np.linspace(0.0, tau, n)
        y = self.q1.copy()
        eta0 = self._defect_from_y(y)
        if not np.isfinite(eta0) or eta0 > self.tol:
            return 0.0
        S = self._step_matrix(tau / (n - 1))
        for k in range(1, n):
            y = S @ y
            eta = self._defect_from_y(y)
            if not np.isfinite(eta) or eta > self.tol:
                lo, hi = float(xs[k - 1]), float(xs[k])
                for _ in range(self.bisect_iters):
                    mid = 0.5 * (lo + hi)
                    if self._defect_from_y(self._reduced_y(mid)) > self.tol:
                        hi = mid
                    else:
                        lo = mid
                return lo
        return float(tau)

    def _reconstruct(self, tau):
        c = self.Z @ self._reduced_y(tau)
        return unvec(self.norm_anchor * (self.V[:, :self.m] @ c), self.d)

    def _record_output(self, rho):
        d = self.diag
        td = float(abs(np.trace(rho) - self.trace_target))
        hd = float(np.linalg.norm(rho - rho.conj().T))
        d["trace_drifts"].append(td)
        d["hermiticity_drifts"].append(hd)
        d["max_trace_drift"] = max(d["max_trace_drift"], td)
        d["max_hermiticity_drift"] = max(d["max_hermiticity_drift"], hd)
        if self.record_min_eig:
            lm = float(np.linalg.eigvalsh((rho + rho.conj().T) / 2)[0])
            d["min_hermitian_eigs"].append(lm)
            d["min_min_hermitian_eig"] = min(d["min_min_hermitian_eig"], lm)

    def _extend(self):
        """_extend_basis! (:460-473)."""
        if not self.extend_basis or self.m >= self.m_max:
            return False
        old = self.m
        self._arnoldi_build(min(self.m + self.extend_step, self.m_max))
        if self.m > old:
            self._build_reduced()
            self.diag["basis_extensions"] += 1
            self.diag["max_dim_used"] = max(self.diag["max_dim_used"], self.m)
            return True
        return False

    def _restart(self, tau):
        """_restart_from_new_anchor! (:475-494)."""
        rho_new = self._reconstruct(tau)
        self.t_anchor += float(tau)
        self.rho_anchor = rho_new
        nrm = np.linalg.norm(rho_new)
        if not nrm > 0:
            raise ValueError("LindbladArnoldiCache: restart failed because anchor density matrix has zero Frobenius norm")
        self.norm_anchor = nrm
        self.Hred[:] = 0
        self.m = 0
        self.V[:, 0] = vec(rho_new) / nrm
        self._arnoldi_build(self.m_init)
        self._build_reduced()
        self.diag["basis_builds"] += 1
        self.diag["restarts"] += 1
        self.diag["max_dim_used"] = max(self.diag["max_dim_used"], self.m)

    def _ensure_valid_up_to(self, target):
        """_ensure_valid_up_to! (:496-535)."""
        rem = float(target)
        if rem <= 0:
            return
        while True:
            v = self._max_valid_interval(rem)
            if v >= rem:
                return
            if self.extend_basis and self.m < self.m_max and self._extend():
                continue
            if v <= 0:
                if self.m < self.m_max and self._extend():
                    continue
                raise ValueError("LindbladArnoldiCache: basis is not valid at tau=0")
            self.diag["accepted_intervals"].append(v)
            self._restart(v)
            rem -= v

    def evolve(self, t):
        """lindblad_timeevolve!(cache, t) (:598-611)."""
        t = float(t)
        if t < self.t_latest:
            raise ValueError("lindblad_timeevolve!: requested time precedes the last served time")
        tau = t - self.t_anchor
        if tau < 0:
            raise ValueError("lindblad_timeevolve!: requested time precedes anchor time")
        self._ensure_valid_up_to(tau)
        rho = self._reconstruct(t - self.t_anchor)
        self._record_output(rho)
        self.t_latest = t
        self.diag["total_times_served"] += 1
        if self.normalize_trace:
            tr = np.trace(rho)
            if tr != 0:
                rho = rho * (self.trace_target / tr)
        return rho


def lindblad_timeevolve(H, jumps, rho0, ts, **kw):
    """lindblad_timeevolve(A, dm, ts; kwargs...) (:537-596): times served in
    sorted order, returned in the given order, with the diagnostics."""
    scalar = np.isscalar(ts)
    ts = np.atleast_1d(np.asarray(ts, float))
    if np.any(ts < 0):
        raise ValueError("lindblad_timeevolve: backward evolution is not supported")
    c = ArnoldiCache(H, jumps, rho0, **kw)
    out = [None] * len(ts)
    for i in np.argsort(ts, kind="stable"):
        out[i] = c.evolve(ts[i])
    return (out[0] if scalar else out), c.diag
