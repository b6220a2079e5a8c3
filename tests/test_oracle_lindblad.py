"""Pins of the CPU Lindblad oracle (oracle/lindblad.py) before it checks the
device: the explicit superoperator against the defining products, the
generator's invariants, closed forms (amplitude damping, the unitary limit,
a steady state), and the adaptive propagator against exp(tL) on the
reference's own test fixtures (test/lindblad_timeevolve_tests.jl:39-330,
their random matrices drawn here from numpy seeds)."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import lindblad as ol


def rand_herm(rng, d):
    A = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    return (A + A.conj().T) / 2


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_superoperator_is_the_defining_products():
    """vec(L(X)) from the Kronecker form equals the product form on random
    non-Hermitian X, and both equal the commutator/anticommutator written out."""
    rng = np.random.default_rng(1)
    d = 4
    H = rand_herm(rng, d)
    J = [rand_c(rng, d, d) / 3 for _ in range(3)]
    X = rand_c(rng, d, d)
    M = ol.explicit_liouvillian(H, J)
    y1 = ol.unvec(M @ ol.vec(X), d)
    y2 = ol.liouvillian_apply(H, J, X)
    y3 = -1j * (H @ X - X @ H)
    for L in J:
        LdL = L.conj().T @ L
        y3 += L @ X @ L.conj().T - 0.5 * (LdL @ X + X @ LdL)
    assert np.abs(y1 - y3).max() < 1e-13
    assert np.abs(y2 - y3).max() < 1e-13
    # a wrong Kronecker order (transposed operand) is caught
    bad = -1j * (np.kron(H, np.eye(d)) - np.kron(np.eye(d), H.T))
    assert np.abs(bad @ ol.vec(X) - ol.vec(-1j * (H @ X - X @ H))).max() > 1e-3


def test_generator_preserves_trace_and_hermiticity():
    """lindblad_timeevolve_tests.jl:192-218."""
    rng = np.random.default_rng(91)
    H = rand_herm(rng, 3)
    J = [rand_c(rng, 3, 3) / 5, rand_c(rng, 3, 3) / 7]
    rh = rand_herm(rng, 3)
    Lr = ol.liouvillian_apply(H, J, rh)
    assert abs(np.trace(Lr)) < 1e-12
    assert np.linalg.norm(Lr - Lr.conj().T) < 1e-12
    for _ in range(3):
        psi = rand_c(rng, 3)
        psi /= np.linalg.norm(psi)
        assert abs(np.trace(ol.liouvillian_apply(H, J, np.outer(psi, psi.conj())))) < 1e-12


def test_amplitude_damping_closed_form():
    """H = w/2 sz, L = sqrt(g) s-: rho_11 decays as exp(-g t), the coherence
    as exp(-(g/2 + i w) t) (|1> the excited state, index 0)."""
    w, g, t = 0.9, 0.35, 1.7
    H = 0.5 * w * np.diag([1.0, -1.0]).astype(complex)
    J = [np.sqrt(g) * np.array([[0, 0], [1, 0]], complex)]
    r0 = np.array([[0.6, 0.3 - 0.2j], [0.3 + 0.2j, 0.4]])
    r = ol.explicit_evolve(H, J, r0, t)
    assert abs(r[0, 0] - 0.6 * np.exp(-g * t)) < 1e-13
    assert abs(r[0, 1] - r0[0, 1] * np.exp(-(g / 2 + 1j * w) * t)) < 1e-13
    rr, _ = ol.lindblad_timeevolve(H, J, r0, t, tol=1e-12, m_init=2, m_max=4)
    assert np.abs(rr - r).max() < 1e-11


def test_single_time_dense_path():
    """lindblad_timeevolve_tests.jl:39-65."""
    H = np.array([[0.2, 0.1 - 0.05j], [0.1 + 0.05j, -0.3]])
    J = [np.sqrt(0.4) * np.array([[0, 1], [0, 0]], complex), np.sqrt(0.15) * np.array([[0, 0], [1, 0]], complex)]
    psi = np.array([0.3 + 0.2j, 0.9])
    r0 = np.outer(psi, psi.conj())
    r, d = ol.lindblad_timeevolve(H, J, r0, 0.7, tol=1e-12, m_init=4, m_max=12, record_min_eig=True)
    assert np.linalg.norm(r - ol.explicit_evolve(H, J, r0, 0.7)) < 1e-10
    assert d["total_times_served"] == 1
    assert len(d["trace_drifts"]) == len(d["hermiticity_drifts"]) == len(d["min_hermitian_eigs"]) == 1
    assert d["max_trace_drift"] == d["trace_drifts"][0]
    assert d["min_min_hermitian_eig"] == d["min_hermitian_eigs"][0]


def test_multi_time_reuse():
    """lindblad_timeevolve_tests.jl:67-112: one basis serves every time."""
    H = np.array([[0.2, 0.1, 0, 0], [0.1, -0.3, 0.15, 0], [0, 0.15, 0.4, 0.12], [0, 0, 0.12, -0.1]], complex)
    J = [np.sqrt(0.2) * np.array([[0, 1, 0, 0], [0, 0, 0, 0], [0, 0, 0, 1], [0, 0, 0, 0]], complex)]
    rng = np.random.default_rng(17)
    psi = rand_c(rng, 4)
    psi /= np.linalg.norm(psi)
    r0 = np.outer(psi, psi.conj())
    ts = np.linspace(0.0, 0.35, 8)
    rs, d = ol.lindblad_timeevolve(H, J, r0, ts, tol=1e-11, m_init=5, m_max=14, record_min_eig=True)
    assert d["total_times_served"] == 8 and d["basis_builds"] == 1 and d["restarts"] == 0
    assert d["max_trace_drift"] == max(d["trace_drifts"])
    for r, t in zip(rs, ts):
        assert np.linalg.norm(r - ol.explicit_evolve(H, J, r0, t)) < 1e-9


def test_unitary_limit_and_restarts():
    """lindblad_timeevolve_tests.jl:138-189: no jumps gives U rho U^H; a
    3-vector basis without extension must restart to reach t = 5."""
    rng = np.random.default_rng(23)
    H = rand_herm(rng, 3)
    psi = rand_c(rng, 3)
    psi /= np.linalg.norm(psi)
    r0 = np.outer(psi, psi.conj())

    def unitary(t):
        U = sla.expm(-1j * t * H)
        return U @ r0 @ U.conj().T

    r, _ = ol.lindblad_timeevolve(H, [], r0, 1.4, tol=1e-12, m_init=4, m_max=9)
    assert np.linalg.norm(r - unitary(1.4)) < 1e-10
    ts = [0.0, 0.2, 0.7, 1.5]
    rs, d = ol.lindblad_timeevolve(H, [], r0, ts, tol=1e-12, m_init=4, m_max=9)
    assert d["total_times_served"] == 4
    for r, t in zip(rs, ts):
        assert np.linalg.norm(r - unitary(t)) < 1e-10
    c = ol.ArnoldiCache(H, [], r0, tol=1e-6, m_init=3, m_max=3, extend_basis=False)
    assert np.linalg.norm(c.evolve(5.0) - unitary(5.0)) < 5e-6
    assert c.diag["restarts"] >= 1


def test_semigroup():
    """lindblad_timeevolve_tests.jl:221-282."""
    rng = np.random.default_rng(202)
    H = rand_herm(rng, 3)
    J = [np.array([[0, 1, 0], [0, 0, 0], [0, 0, 0]], complex) * np.sqrt(0.2),
         np.array([[0, 0, 0], [0, 0, 1], [0, 0, 0]], complex) * np.sqrt(0.15)]
    psi = rand_c(rng, 3)
    r0 = np.outer(psi, psi.conj())
    kw = dict(tol=1e-11, m_init=4, m_max=9)
    one, _ = ol.lindblad_timeevolve(H, J, r0, 0.85, **kw)
    a, _ = ol.lindblad_timeevolve(H, J, r0, 0.25, **kw)
    seq, _ = ol.lindblad_timeevolve(H, J, a, 0.6, **kw)
    assert np.linalg.norm(one - seq) < 1e-9
    c = ol.ArnoldiCache(H, J, r0, **kw)
    c.evolve(0.25)
    assert np.linalg.norm(one - c.evolve(0.85)) < 1e-9
    kr = dict(tol=1e-6, m_init=5, m_max=5, extend_basis=False)
    c = ol.ArnoldiCache(H, J, r0, **kr)
    c.evolve(0.6)
    st = c.evolve(2.0)
    tot, _ = ol.lindblad_timeevolve(H, J, r0, 2.0, **kr)
    assert np.linalg.norm(st - tot) < 1e-6 and c.diag["restarts"] >= 1
    assert np.linalg.norm(st - ol.explicit_evolve(H, J, r0, 2.0)) < 1e-5


def test_steady_state_and_forward_only():
    """lindblad_timeevolve_tests.jl:122-135 and 284-305."""
    J = [np.sqrt(0.4) * np.array([[0, 1], [0, 0]], complex)]
    rss = np.diag([1.0, 0.0]).astype(complex)
    rs, d = ol.lindblad_timeevolve(np.zeros((2, 2)), J, rss, np.linspace(0, 3, 9), tol=1e-12, m_init=2, m_max=4,
                                   record_min_eig=True)
    for r in rs:
        assert np.linalg.norm(r - rss) < 1e-12
    assert max(d["trace_drifts"]) < 1e-12 and d["min_min_hermitian_eig"] >= -1e-12
    H = np.array([[0, 1], [1, 0]], complex)
    with pytest.raises(ValueError):
        ol.lindblad_timeevolve(H, J, rss, -0.1)
    c = ol.ArnoldiCache(H, J, rss, tol=1e-10, m_init=4, m_max=10)
    c.evolve(0.2)
    with pytest.raises(ValueError):
        c.evolve(0.1)


def test_tightening_controls():
    """lindblad_timeevolve_tests.jl:333-371."""
    rng = np.random.default_rng(66)
    H = rand_herm(rng, 4)
    J = [np.array([[0, 1, 0, 0], [0, 0, 0, 0], [0, 0, 0, 1], [0, 0, 0, 0]], complex) * np.sqrt(0.2),
         rand_c(rng, 4, 4) / 10]
    psi = rand_c(rng, 4)
    r0 = np.outer(psi, psi.conj())
    ts = [0.4, 1.0, 1.7]
    lo, _ = ol.lindblad_timeevolve(H, J, r0, ts, tol=1e-2, m_init=2, m_max=2, extend_basis=False)
    hi, _ = ol.lindblad_timeevolve(H, J, r0, ts, tol=1e-12, m_init=6, m_max=12)
    e_lo = max(np.linalg.norm(r - ol.explicit_evolve(H, J, r0, t)) for r, t in zip(lo, ts))
    e_hi = max(np.linalg.norm(r - ol.explicit_evolve(H, J, r0, t)) for r, t in zip(hi, ts))
    assert e_hi < e_lo and e_hi < 1e-9


def test_dissipative_chain_against_expm():
    """A 4-site dissipative XXZ chain (d = 16, one decay jump per site):
    the propagator against exp(tL) with restarts and extensions in play."""
    from tests.lindblad_models import dissipative_xxz
    H, J = dissipative_xxz(4, 1.0, 0.5, 0.2)
    rng = np.random.default_rng(5)
    psi = rand_c(rng, 16)
    psi /= np.linalg.norm(psi)
    r0 = np.outer(psi, psi.conj())
    ts = [0.3, 1.1, 2.5]
    rs, d = ol.lindblad_timeevolve(H, J, r0, ts, tol=1e-10, m_init=8, m_max=16)
    assert d["restarts"] >= 1
    for r, t in zip(rs, ts):
        assert np.linalg.norm(r - ol.explicit_evolve(H, J, r0, t)) < 1e-8
