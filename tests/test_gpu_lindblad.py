"""GPU parity of the Lindblad propagator (LiouvillianMap, LindbladArnoldiCache,
lindblad_timeevolve; LindbladEvolution.jl:62-611): the device generator
against the oracle's products, the device propagator against the oracle's
restatement of the same controller (same counters, same accepted intervals)
and against exp(tL), on the reference's own test fixtures
(test/lindblad_timeevolve_tests.jl) and on dissipative XXZ chains up to
d = 256."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import lindblad as ol
from tests.lindblad_models import dissipative_xxz

pytestmark = pytest.mark.gpu


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rand_herm(rng, d):
    A = rand_c(rng, d, d)
    return (A + A.conj().T) / 2


@pytest.mark.parametrize("d,nj", [(1, 0), (2, 0), (3, 2), (4, 1), (16, 4), (64, 6), (100, 3)])
def test_liouvillian_apply_matches_oracle(gpu, d, nj):
    """Random complex (also non-Hermitian) H, jumps and X: FP64 products
    agree to rounding; the vec(rho) form gives the same numbers."""
    import edkit_hip as edk
    rng = np.random.default_rng(100 + d + nj)
    H = rand_c(rng, d, d)
    J = [rand_c(rng, d, d) / np.sqrt(d) for _ in range(nj)]
    X = rand_c(rng, d, d)
    A = edk.LiouvillianMap(H, J)
    Y = A @ X
    Yo = ol.liouvillian_apply(H, J, X)
    assert np.abs(Y - Yo).max() <= 1e-12 * max(1.0, np.abs(Yo).max())
    yv = A @ ol.vec(X)
    assert np.abs(ol.unvec(yv, d) - Y).max() == 0.0
    if d <= 16:
        M = ol.explicit_liouvillian(H, J)
        assert np.abs(ol.vec(Y) - M @ ol.vec(X)).max() <= 1e-12 * max(1.0, np.abs(Yo).max())


def fixture(name):
    """The reference's test fixtures (lindblad_timeevolve_tests.jl), random
    parts from numpy seeds."""
    rng = np.random.default_rng({"single": 1, "multi": 17, "unitary": 23, "semigroup": 202, "steady": 3,
                                 "parity": 44, "tight": 66}[name])
    if name == "single":
        H = np.array([[0.2, 0.1 - 0.05j], [0.1 + 0.05j, -0.3]])
        J = [np.sqrt(0.4) * np.array([[0, 1], [0, 0]], complex), np.sqrt(0.15) * np.array([[0, 0], [1, 0]], complex)]
        psi = np.array([0.3 + 0.2j, 0.9])
        return H, J, np.outer(psi, psi.conj()), [0.7], dict(tol=1e-12, m_init=4, m_max=12, record_min_eig=True)
    if name == "multi":
        H = np.array([[0.2, 0.1, 0, 0], [0.1, -0.3, 0.15, 0], [0, 0.15, 0.4, 0.12], [0, 0, 0.12, -0.1]], complex)
        J = [np.sqrt(0.2) * np.array([[0, 1, 0, 0], [0, 0, 0, 0], [0, 0, 0, 1], [0, 0, 0, 0]], complex)]
        psi = rand_c(rng, 4)
        psi /= np.linalg.norm(psi)
        return (H, J, np.outer(psi, psi.conj()), list(np.linspace(0, 0.35, 8)),
                dict(tol=1e-11, m_init=5, m_max=14, record_min_eig=True))
    if name == "unitary":
        H = rand_herm(rng, 3)
        psi = rand_c(rng, 3)
        psi /= np.linalg.norm(psi)
        return H, [], np.outer(psi, psi.conj()), [0.0, 0.2, 0.7, 1.5], dict(tol=1e-12, m_init=4, m_max=9)
    if name == "semigroup":
        H = rand_herm(rng, 3)
        J = [np.array([[0, 1, 0], [0, 0, 0], [0, 0, 0]], complex) * np.sqrt(0.2),
             np.array([[0, 0, 0], [0, 0, 1], [0, 0, 0]], complex) * np.sqrt(0.15)]
        psi = rand_c(rng, 3)
        return H, J, np.outer(psi, psi.conj()), [0.6, 2.0], dict(tol=1e-6, m_init=5, m_max=5, extend_basis=False)
    if name == "steady":
        J = [np.sqrt(0.4) * np.array([[0, 1], [0, 0]], complex)]
        return (np.zeros((2, 2), complex), J, np.diag([1.0, 0.0]).astype(complex), list(np.linspace(0, 3, 9)),
                dict(tol=1e-12, m_init=2, m_max=4, record_min_eig=True))
    if name == "parity":
        H = rand_herm(rng, 4)
        J = [rand_c(rng, 4, 4) / 6, rand_c(rng, 4, 4) / 8]
        psi = rand_c(rng, 4)
        return (H, J, np.outer(psi, psi.conj()), [0.0, 0.15, 0.45, 0.8],
                dict(tol=1e-11, m_init=5, m_max=12, record_min_eig=True))
    H = rand_herm(rng, 4)
    J = [np.array([[0, 1, 0, 0], [0, 0, 0, 0], [0, 0, 0, 1], [0, 0, 0, 0]], complex) * np.sqrt(0.2),
         rand_c(rng, 4, 4) / 10]
    psi = rand_c(rng, 4)
    return H, J, np.outer(psi, psi.conj()), [0.4, 1.0, 1.7], dict(tol=1e-2, m_init=2, m_max=2, extend_basis=False)


COUNTERS = ("basis_builds", "basis_extensions", "restarts", "liouvillian_applies", "total_times_served",
            "max_dim_used")


def compare_runs(rs, d, ro, do):
    """Counters equal; accepted intervals to 1e-3 relative (each ends where
    the defect, a component of exp(tau Hm) e1 near the tolerance, crosses the
    tolerance, and that component carries only a few correct digits); the
    density matrices to 1e-12 without restarts, 1e-9 across restarts at
    such anchors."""
    for k in COUNTERS:
        assert d[k] == do[k], k
    assert len(d["accepted_intervals"]) == len(do["accepted_intervals"])
    assert np.allclose(d["accepted_intervals"], do["accepted_intervals"], rtol=1e-3, atol=0)
    tol = 1e-12 if do["restarts"] == 0 else 1e-9
    for r, q in zip(rs, ro):
        assert np.abs(r - q).max() <= tol
    for key in ("trace_drifts", "hermiticity_drifts", "min_hermitian_eigs"):
        assert len(d[key]) == len(do[key])
        assert np.allclose(d[key], do[key], rtol=0, atol=max(tol, 1e-13)), key


@pytest.mark.parametrize("name", ["single", "multi", "unitary", "semigroup", "steady", "parity", "tight"])
def test_reference_fixtures_match_oracle(gpu, name):
    """Same counters and accepted intervals as the oracle, the same density
    matrices, and exp(tL) within the reference's bounds."""
    import edkit_hip as edk
    H, J, r0, ts, kw = fixture(name)
    out, diag = edk.lindblad_timeevolve(edk.lindblad(H, J), r0, ts, return_diagnostics=True, **kw)
    rs = [o.rho for o in out]
    ro, do = ol.lindblad_timeevolve(H, J, r0, ts, **kw)
    compare_runs(rs, diag, ro, do)
    bound = {1e-12: 1e-10, 1e-11: 1e-9}.get(kw["tol"])
    if bound:
        for r, t in zip(rs, ts):
            assert np.linalg.norm(r - ol.explicit_evolve(H, J, r0, t)) < bound
    if name == "semigroup":
        assert diag["restarts"] >= 1
        assert np.linalg.norm(rs[-1] - ol.explicit_evolve(H, J, r0, 2.0)) < 1e-5
    if name == "steady":
        assert max(diag["trace_drifts"]) < 1e-12 and diag["min_min_hermitian_eig"] >= -1e-12


def test_cache_workflow_and_forward_only(gpu):
    """lindblad_timeevolve_tests.jl:114-150: a LiouvillianMap cache served
    step by step, and the forward-only errors."""
    import edkit_hip as edk
    H = np.array([[0.0, 0.2], [0.2, 0.1]], complex)
    J = [np.sqrt(0.3) * np.array([[0, 1], [0, 0]], complex)]
    A = edk.LiouvillianMap(H, J)
    c = edk.LindbladArnoldiCache(A, edk.densitymatrix(np.array([0.0, 1.0])), tol=1e-12, m_init=4, m_max=10,
                                 record_min_eig=True)
    r = edk.lindblad_timeevolve_(c, 0.5)
    r0 = np.diag([0.0, 1.0]).astype(complex)
    assert np.linalg.norm(r.rho - ol.explicit_evolve(H, J, r0, 0.5)) < 1e-10
    dg = c.diagnostics
    assert dg["total_times_served"] == 1 and dg["liouvillian_applies"] >= c.m and len(dg["trace_drifts"]) == 1
    with pytest.raises(edk.EdkError, match="precedes the last served time"):
        c.evolve(0.1)
    with pytest.raises(ValueError):
        edk.lindblad_timeevolve(A, r0, -0.1)
    with pytest.raises(ValueError):
        edk.lindblad_timeevolve(A, r0, [-0.1, 0.2])


def test_argument_errors(gpu):
    import edkit_hip as edk
    A = edk.LiouvillianMap(np.eye(2), [])
    r0 = np.diag([1.0, 0.0])
    for kw, msg in ((dict(m_init=0), "m_init"), (dict(m_init=5, m_max=4), "m_max"), (dict(extend_step=0), "extend_step"),
                    (dict(nsample=1), "nsample"), (dict(bisect_iters=-1), "bisect_iters")):
        with pytest.raises(edk.EdkError, match=msg):
            edk.LindbladArnoldiCache(A, r0, **kw)
    with pytest.raises(edk.EdkError, match="zero Frobenius norm"):
        edk.LindbladArnoldiCache(A, np.zeros((2, 2)))
    with pytest.raises(ValueError, match="incompatible size"):
        edk.LindbladArnoldiCache(A, np.eye(3))
    with pytest.raises(ValueError, match="jump 1"):
        edk.LiouvillianMap(np.eye(2), [np.eye(3)])


def test_dissipative_chain_d32_against_expm(gpu):
    """5-site dissipative XXZ chain (d = 32, five decay jumps): the device
    against exp(tL) (a 1024 x 1024 exponential) and against the oracle's
    counters, with extensions and restarts in play."""
    import edkit_hip as edk
    H, J = dissipative_xxz(5, 1.0, 0.5, 0.2)
    rng = np.random.default_rng(7)
    psi = rand_c(rng, 32)
    psi /= np.linalg.norm(psi)
    r0 = np.outer(psi, psi.conj())
    ts = [0.5, 2.0, 4.0]
    kw = dict(tol=1e-10, m_init=10, m_max=24, record_min_eig=True)
    out, diag = edk.lindblad_timeevolve(edk.lindblad(H, J), r0, ts, return_diagnostics=True, **kw)
    ro, do = ol.lindblad_timeevolve(H, J, r0, ts, **kw)
    compare_runs([o.rho for o in out], diag, ro, do)
    assert diag["restarts"] >= 1
    M = ol.explicit_liouvillian(H, J)
    for o, t in zip(out, ts):
        ref = ol.unvec(sla.expm(t * M) @ ol.vec(r0), 32)
        assert np.linalg.norm(o.rho - ref) < 1e-8
        assert abs(np.trace(o.rho) - 1) < 1e-9


def test_chain_d256_invariants(gpu):
    """8-site chain (d = 256, eight jumps; 65 536-dimensional Liouville
    space): trace and Hermiticity kept, no negative eigenvalue beyond the
    tolerance, the semigroup property across a cache, and the unitary limit
    against U rho U^H."""
    import edkit_hip as edk
    H, J = dissipative_xxz(8, 1.0, 0.5, 0.1)
    rng = np.random.default_rng(8)
    psi = rand_c(rng, 256)
    psi /= np.linalg.norm(psi)
    r0 = np.outer(psi, psi.conj())
    kw = dict(tol=1e-10, m_init=20, m_max=40, record_min_eig=True)
    c = edk.LindbladArnoldiCache(edk.lindblad(H, J), r0, **kw)
    c.evolve(0.5)
    r_c = c.evolve(1.5).rho
    dg = c.diagnostics
    assert max(dg["trace_drifts"]) < 1e-9 and max(dg["hermiticity_drifts"]) < 1e-9
    assert dg["min_min_hermitian_eig"] > -1e-9
    r_one = edk.lindblad_timeevolve(edk.lindblad(H, J), r0, 1.5, **kw).rho
    assert np.linalg.norm(r_c - r_one) < 1e-8
    U = sla.expm(-1j * 0.8 * H)
    r_u = edk.lindblad_timeevolve(edk.lindblad(H, []), r0, 0.8, **kw).rho
    assert np.linalg.norm(r_u - U @ r0 @ U.conj().T) < 1e-9


def test_normalize_trace(gpu):
    """normalize_trace rescales each output to the initial trace
    (_copy_output_density, LindbladEvolution.jl:451-458) after the drift is
    recorded on the raw output; a loose tolerance leaves a visible drift."""
    import edkit_hip as edk
    H, J = dissipative_xxz(3, 1.0, 0.5, 0.3)
    r0 = np.diag(np.linspace(1, 2, 8)).astype(complex)
    r0 /= np.trace(r0)
    kw = dict(tol=1e-3, m_init=2, m_max=3, extend_basis=False)
    raw, d_raw = edk.lindblad_timeevolve(edk.lindblad(H, J), r0, 3.0, return_diagnostics=True, **kw)
    nrm, d_n = edk.lindblad_timeevolve(edk.lindblad(H, J), r0, 3.0, return_diagnostics=True, normalize_trace=True,
                                       **kw)
    assert np.allclose(d_raw["trace_drifts"], d_n["trace_drifts"], rtol=1e-12, atol=0)
    assert abs(np.trace(nrm.rho) - 1) < 1e-14
    assert np.abs(nrm.rho - raw.rho / np.trace(raw.rho)).max() < 1e-14
    ro, _ = ol.lindblad_timeevolve(H, J, r0, 3.0, normalize_trace=True, **kw)
    assert np.abs(nrm.rho - ro).max() < 1e-9


def rand_sparse(rng, d, density):
    """Random complex CSR matrix with round(density d^2) (at least one) distinct entries."""
    import scipy.sparse as sp
    nnz = max(1, int(round(density * d * d)))
    idx = rng.choice(d * d, size=nnz, replace=False)
    vals = rng.standard_normal(nnz) + 1j * rng.standard_normal(nnz)
    return sp.csr_matrix((vals, (idx % d, idx // d)), shape=(d, d))


@pytest.mark.parametrize("d,nj,density", [(1, 1, 1.0), (2, 2, 0.6), (4, 1, 0.5), (16, 3, 0.2), (64, 5, 0.05),
                                          (256, 4, 0.01)])
def test_sparse_apply_matches_oracle_and_dense(gpu, d, nj, density):
    """The sparse form (CSR row and column passes over rho) against the
    oracle's products and against the dense form, non-Hermitian H included."""
    import edkit_hip as edk
    rng = np.random.default_rng(300 + d)
    H = rand_sparse(rng, d, density)
    J = [rand_sparse(rng, d, density) for _ in range(nj)]
    X = rand_c(rng, d, d)
    As = edk.LiouvillianMap(H, J)
    assert As.sparse
    Ad = edk.LiouvillianMap(H, J, sparse=False)
    Ys = As @ X
    Yo = ol.liouvillian_apply(H.toarray(), [L.toarray() for L in J], X)
    assert np.abs(Ys - Yo).max() <= 1e-12 * max(1.0, np.abs(Yo).max())
    assert np.abs(Ys - Ad @ X).max() <= 1e-12 * max(1.0, np.abs(Yo).max())


def test_sparse_cache_workflow(gpu):
    """lindblad_timeevolve_tests.jl:114-135: a sparse LiouvillianMap cache."""
    import edkit_hip as edk
    import scipy.sparse as sp
    H = sp.csr_matrix(np.array([[0.0, 0.2], [0.2, 0.1]], complex))
    J = [sp.csr_matrix(np.sqrt(0.3) * np.array([[0, 1], [0, 0]], complex))]
    A = edk.LiouvillianMap(H, J)
    c = edk.LindbladArnoldiCache(A, edk.densitymatrix(np.array([0.0, 1.0])), tol=1e-12, m_init=4, m_max=10,
                                 record_min_eig=True)
    r = edk.lindblad_timeevolve_(c, 0.5)
    ref = ol.explicit_evolve(H.toarray(), [J[0].toarray()], np.diag([0.0, 1.0]).astype(complex), 0.5)
    assert np.linalg.norm(r.rho - ref) < 1e-10
    dg = c.diagnostics
    assert dg["total_times_served"] == 1 and dg["liouvillian_applies"] >= c.m and len(dg["trace_drifts"]) == 1


def test_dense_and_sparse_backend_parity(gpu):
    """lindblad_timeevolve_tests.jl:307-331: the legacy object, the dense and
    the sparse map give the same states (1e-10) and diagnostics (1e-12)."""
    import edkit_hip as edk
    import scipy.sparse as sp
    H, J, r0, ts, kw = fixture("parity")
    out = {}
    for name, A in (("lb", edk.lindblad(H, J)), ("dense", edk.LiouvillianMap(H, J)),
                    ("sparse", edk.LiouvillianMap(sp.csr_matrix(H), [sp.csr_matrix(L) for L in J]))):
        out[name] = edk.lindblad_timeevolve(A, r0, ts, return_diagnostics=True, **kw)
    for name in ("dense", "sparse"):
        for a, b in zip(out["lb"][0], out[name][0]):
            assert np.linalg.norm(a.rho - b.rho) < 1e-10
        for key in ("trace_drifts", "hermiticity_drifts", "min_hermitian_eigs"):
            assert np.allclose(out["lb"][1][key], out[name][1][key], rtol=0, atol=1e-12), (name, key)


def test_sparse_chain_against_dense(gpu):
    """8-site dissipative chain (d = 256) through the sparse and the dense
    forms: same counters, states within 1e-9."""
    import edkit_hip as edk
    import scipy.sparse as sp
    H, J = dissipative_xxz(8, 1.0, 0.5, 0.1)
    rng = np.random.default_rng(12)
    psi = rand_c(rng, 256)
    psi /= np.linalg.norm(psi)
    r0 = np.outer(psi, psi.conj())
    kw = dict(tol=1e-10, m_init=20, m_max=40)
    ts = [0.5, 1.5, 3.0]
    rd, dd = edk.lindblad_timeevolve(edk.LiouvillianMap(H, J), r0, ts, return_diagnostics=True, **kw)
    rs, ds = edk.lindblad_timeevolve(edk.LiouvillianMap(sp.csr_matrix(H), [sp.csr_matrix(L) for L in J]), r0, ts,
                                     return_diagnostics=True, **kw)
    for k in COUNTERS:
        assert dd[k] == ds[k], k
    for a, b in zip(rd, rs):
        assert np.abs(a.rho - b.rho).max() < 1e-9


def test_sparse_chain_d4096_invariants(gpu):
    """12-site dissipative chain (d = 4096, Liouville dimension 16.8 M) on the
    sparse form only: trace and Hermiticity kept, and the semigroup property
    across a cache."""
    import edkit_hip as edk
    import scipy.sparse as sp
    from tests.lindblad_models import SX, SY, SZ

    def site(op, i, L):
        return sp.kron(sp.kron(sp.identity(2 ** i), sp.csr_matrix(op)), sp.identity(2 ** (L - i - 1)), format="csr")

    L = 12
    H = sp.csr_matrix((2 ** L, 2 ** L), dtype=complex)
    for i in range(L - 1):
        for P, c in ((SX, 1.0), (SY, 1.0), (SZ, 0.5)):
            H = H + 0.25 * c * site(P, i, L) @ site(P, i + 1, L)
    sm = np.array([[0, 0], [1, 0]], complex)
    J = [np.sqrt(0.1) * site(sm, i, L) for i in range(L)]
    rng = np.random.default_rng(13)
    psi = rand_c(rng, 2 ** L)
    psi /= np.linalg.norm(psi)
    r0 = np.outer(psi, psi.conj())
    A = edk.LiouvillianMap(H, J)
    c = edk.LindbladArnoldiCache(A, r0, tol=1e-10, m_init=20, m_max=40)
    c.evolve(0.3)
    r_c = c.evolve(0.8).rho
    dg = c.diagnostics
    assert max(dg["trace_drifts"]) < 1e-9 and max(dg["hermiticity_drifts"]) < 1e-9
    r_one = edk.lindblad_timeevolve(A, r0, 0.8, tol=1e-10, m_init=20, m_max=40).rho
    assert np.abs(r_c - r_one).max() < 1e-9
