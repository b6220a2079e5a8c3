"""Pins of the CPU DoubleBasis oracle (oracle/linearmap.py): embeddings with
orthonormal columns, the reference's own DoubleBasis tests
(test/advanced_tests.jl:1-60, test/DoubleBasisTest.jl), and the rectangular
operator against the compression Q1^H O Q2 (Q = conj(embedding)) of the
full-space operator, a different computation from the column scatter."""
import numpy as np
import pytest

from oracle import linearmap as olm
from oracle import oracle as orc

SP = np.array([[0, 1], [0, 0]], complex)

KINDS4 = [("tensor", {}), ("projected", dict(N=2)), ("parity", dict(p=1)), ("flip", dict(p=1, N=2)),
          ("parity_flip", dict(p=1, z=1, N=2)), ("translational", dict(k=0, N=2)),
          ("translation_parity", dict(k=0, p=1, N=2)), ("translation_flip", dict(k=0, p=1, N=2))]


def ob_of(kind, L, kw, base=2):
    if kind == "abelian":
        return orc.Basis("abelian", L, base=base, N=kw.get("N"), k=kw.get("k", 0), p=kw.get("p", 0),
                         z=kw.get("z", 0), has_k="k" in kw, has_p="p" in kw, has_z="z" in kw)
    return orc.Basis(kind, L, base=base, **kw)


@pytest.mark.parametrize("kind,kw", KINDS4 + [("translational", dict(k=1, N=3)),
                                               ("translation_flip", dict(k=2, p=-1, N=3))])
def test_embedding_columns_orthonormal(kind, kw):
    L = 4 if kw.get("N", 2) == 2 else 6
    E = olm.basis_embedding(ob_of(kind, L, kw))
    assert np.abs(E.conj().T @ E - np.eye(E.shape[1])).max() < 1e-14


def test_reference_linear_maps():
    """advanced_tests.jl:1-45: the identity symmetrizer, the parity
    embedding, and T(v) = symmetrizer(T) v for every pair of basis types."""
    L = 4
    full = orc.Basis("tensor", L)
    rng = np.random.default_rng(4)
    psi = rng.standard_normal(16) + 1j * rng.standard_normal(16)
    assert np.abs(olm.symmetrizer(full, full) - np.eye(16)).max() < 1e-15
    assert np.abs(olm.double_map(full, full, psi) - psi).max() < 1e-15
    bpar = orc.Basis("parity", L, p=1)
    S = olm.symmetrizer(bpar, full)
    assert np.abs(S - olm.basis_embedding(bpar).conj().T).max() < 1e-15
    P = olm.basis_embedding(bpar)
    assert np.abs(P @ (S @ psi) - P @ (P.conj().T @ psi)).max() < 1e-14
    bases = KINDS4 + [("abelian", dict(N=2, k=0, p=1))]
    for kt, kwt in bases:
        for ks, kws in bases:
            bt, bs = ob_of(kt, L, kwt), ob_of(ks, L, kws)
            v = rng.standard_normal(bs.D) + 1j * rng.standard_normal(bs.D)
            assert np.abs(olm.double_map(bt, bs, v) - olm.symmetrizer(bt, bs) @ v).max() < 1e-13, (kt, ks)


def test_rectangular_operator_compressions():
    """S+ from N to N+1 (advanced_tests.jl:50-58) and covariant operators
    between momentum and reflection sectors: Q1^H O Q2 with Q = conj(P)."""
    L = 6
    full = orc.Basis("tensor", L)
    one = [[i] for i in range(1, L + 1)]
    O = olm.double_operator(full, full, [SP] * L, one)
    b1, b2 = orc.Basis("projected", L, N=3), orc.Basis("projected", L, N=2)
    H = olm.double_operator(b1, b2, [SP] * L, one)
    assert H.shape == (b1.D, b2.D)
    assert np.abs(H - olm.basis_embedding(b1).T @ O @ olm.basis_embedding(b2)).max() < 1e-14
    for k2 in range(L):
        for q in (1, 2, 3):
            b1 = orc.Basis("translational", L, N=3, k=(k2 + q) % L)
            b2 = orc.Basis("translational", L, N=2, k=k2)
            mats = [np.exp(2j * np.pi * q * i / L) * SP for i in range(1, L + 1)]
            H = olm.double_operator(b1, b2, mats, one)
            Of = olm.double_operator(full, full, mats, one)
            Q1, Q2 = olm.basis_embedding(b1).conj(), olm.basis_embedding(b2).conj()
            assert H.size == 0 or np.abs(H - Q1.conj().T @ Of @ Q2).max() < 1e-13, (k2, q)
            # the other convention (P itself) relabels k -> -k and fails
            if q in (1, 2) and b1.D and b2.D:
                P1, P2 = olm.basis_embedding(b1), olm.basis_embedding(b2)
                assert np.abs(H - P1.conj().T @ Of @ P2).max() > 1e-3
    for p in (1, -1):
        b1 = orc.Basis("translation_parity", L, N=3, k=0, p=p)
        b2 = orc.Basis("translation_parity", L, N=2, k=0, p=p)
        H = olm.double_operator(b1, b2, [SP] * L, one)
        Q1, Q2 = olm.basis_embedding(b1).conj(), olm.basis_embedding(b2).conj()
        assert H.size == 0 or np.abs(H - Q1.conj().T @ O @ Q2).max() < 1e-13
        assert H.shape == (b1.D, b2.D)


def test_spin1_transition_weights():
    """DoubleBasisTest.jl:15-46: the summed transition weights |<e1|S|e2>|^2
    over all pairs of neighbouring magnetisation sectors equal the full-space
    sum, for a random raising-type operator on three spin-1 sites."""
    L = 4
    rng = np.random.default_rng(15)
    p3, p2, p1, z0 = [1], [2, 4, 10], [3, 5, 7, 11, 13, 19], [6, 8, 12, 14, 16, 20, 22]
    m1, m2, m3 = [9, 15, 17, 21, 23, 25], [18, 24, 26], [27]
    Sp = np.zeros((27, 27), complex)
    for a, b in ((m2, m3), (m1, m2), (z0, m1), (p1, z0), (p2, p1), (p3, p2)):
        blk = rng.random((len(a), len(b))) + 1j * rng.random((len(a), len(b)))
        Sp[np.ix_(np.array(a) - 1, np.array(b) - 1)] = blk
    s1 = np.array([[1, 0, 0], [0, 0, 0], [0, 0, -1]], complex)
    spl = np.sqrt(2) * np.array([[0, 1, 0], [0, 0, 1], [0, 0, 0]], complex)
    XY = np.kron(spl, spl.conj().T) + np.kron(spl.conj().T, spl) + np.kron(s1, np.eye(3)) + np.kron(np.eye(3), s1)
    bonds = [[i, i % L + 1] for i in range(1, L + 1)]
    trip = [[i, (i % L) + 1, ((i + 1) % L) + 1] for i in range(1, L + 1)]
    full = orc.Basis("tensor", L, base=3)
    Hf = olm.double_operator(full, full, [XY] * L, bonds)
    Of = olm.double_operator(full, full, [Sp] * L, trip)
    _, v = np.linalg.eigh(0.5 * (Hf + Hf.conj().T))
    B = np.sum(np.abs(v.conj().T @ Of @ v) ** 2)
    A = 0.0
    # ProjectedBasis(f = x -> sum(x) == n) on the digits: N counts
    # sum(base - 1 - digit), so digit sum n is N = 2L - n
    for n in range(1, 2 * L + 1):
        b1 = orc.Basis("projected", L, base=3, N=2 * L - (n - 1))
        b2 = orc.Basis("projected", L, base=3, N=2 * L - n)
        h1 = olm.double_operator(b1, b1, [XY] * L, bonds)
        h2 = olm.double_operator(b2, b2, [XY] * L, bonds)
        _, v1 = np.linalg.eigh(0.5 * (h1 + h1.conj().T))
        _, v2 = np.linalg.eigh(0.5 * (h2 + h2.conj().T))
        op = olm.double_operator(b1, b2, [Sp] * L, trip)
        A += np.sum(np.abs(v1.conj().T @ op @ v2) ** 2)
    assert abs(A - B) < 1e-8 * B
    # every block is needed: dropping one sector pair breaks the sum
    assert A - np.sum(np.abs(op) ** 2) < B * (1 - 1e-6)
