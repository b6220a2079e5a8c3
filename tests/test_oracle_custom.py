"""Custom permutation generators (basis(...; symmetries), AbelianBasis.jl:771-847,
the 2D example at :754-760): the oracle's tables against the definition,
its sector dimensions against the charge sector, and the spectra of all
momentum sectors of a 2D torus against the charge sector's spectrum."""
import itertools

import numpy as np
import pytest

from oracle import oracle as orc
from tests import dense
from tests.models import Shape, heisenberg_2d, torus_translations


def gens_of(perm, q, order, inv=None):
    """dense.group_elements generator for AbelianOperator(order, q, perm):
    character exp(2 pi i q m / order) for the m-th power."""
    return (perm, inv, order, [np.exp(2j * np.pi * q * m / order) for m in range(order)])


@pytest.mark.parametrize("Lx,Ly,N", [(4, 2, 4), (3, 3, 4), (3, 3, 5)])
def test_torus_tables_by_definition(Lx, Ly, N):
    L = Lx * Ly
    tx, ty = torus_translations(Lx, Ly)
    total = 0
    for kx, ky in itertools.product(range(Lx), range(Ly)):
        ob = orc.Basis("abelian", L, N=N, symmetries=[(tx, kx), (ty, ky)])
        I, R = ob.tables()
        # base 2 with N takes the Gosper path, states with N set bits
        # (AbelianBasis.jl:471-485, 530-535), not its predicate sum(x) == L - N
        reps, norms = dense.sector_by_definition(L, 2, [gens_of(tx, kx, Lx), gens_of(ty, ky, Ly)],
                                                 lambda d: sum(d) == N)
        assert np.array_equal(I - 1, reps), (kx, ky)
        np.testing.assert_allclose(R, norms, rtol=1e-13)
        total += ob.D
    # every orbit is counted once per compatible momentum: the momentum sectors
    # partition the charge sector
    assert total == orc.Basis("projected", L, N=N).D


def test_custom_flip_generator_with_inversion():
    """A spin flip given as a custom symmetry (identity permutation, every site
    inverted) beside T_x, T_y of a 4 x 2 torus at half filling."""
    Lx, Ly, L, N = 4, 2, 8, 4
    tx, ty = torus_translations(Lx, Ly)
    ident, allinv = list(range(1, L + 1)), [1] * L
    for kx, ky, q in itertools.product(range(Lx), range(Ly), (0, 1)):
        ob = orc.Basis("abelian", L, N=N, symmetries=[(tx, kx), (ty, ky), (ident, q, allinv)])
        I, R = ob.tables()
        reps, norms = dense.sector_by_definition(
            L, 2, [gens_of(tx, kx, Lx), gens_of(ty, ky, Ly), gens_of(ident, q, 2, allinv)],
            lambda d: sum(d) == N)
        assert np.array_equal(I - 1, reps), (kx, ky, q)
        np.testing.assert_allclose(R, norms, rtol=1e-13)


def _sector_matrix(op, D):
    return np.stack([op @ np.eye(D, dtype=np.complex128)[:, j] for j in range(D)], axis=1)


def test_torus_spectrum_union_over_momenta():
    """4 x 3 torus, XXZ (Delta = 0.6), N = 6: the 12 momentum sectors
    reassemble the 924-state charge sector's spectrum."""
    Lx, Ly, N = 4, 3, 6
    L = Lx * Ly
    H = heisenberg_2d(Shape(L), Lx, Ly, delta=0.6)
    O = dense.full_operator(H, L)
    I, _ = orc.Basis("projected", L, N=N).tables()
    full = np.sort(np.linalg.eigvalsh(O[np.ix_(I - 1, I - 1)]))
    tx, ty = torus_translations(Lx, Ly)
    vals = []
    for kx, ky in itertools.product(range(Lx), range(Ly)):
        ob = orc.Basis("abelian", L, N=N, symmetries=[(tx, kx), (ty, ky)])
        Hs = _sector_matrix(orc.Op(ob, H.flatten()), ob.D)
        np.testing.assert_allclose(Hs, Hs.conj().T, atol=1e-12)
        vals += list(np.linalg.eigvalsh(Hs))
    np.testing.assert_allclose(np.sort(vals), full, atol=1e-11)


@pytest.mark.parametrize("kx,ky", [(0, 0), (1, 2), (3, 1)])
def test_torus_schmidt_is_the_embedded_state(kx, ky):
    """schmidt(v, A, b) for a custom-generator basis (4 x 3 torus, N = 6)
    element by element against the full-space state conj(P) v, P the
    reference's embedding coef / order(b), reshaped by the A and B digits."""
    Lx, Ly, N = 4, 3, 6
    L = Lx * Ly
    tx, ty = torus_translations(Lx, Ly)
    ob = orc.Basis("abelian", L, N=N, symmetries=[(tx, kx), (ty, ky)])
    rng = np.random.default_rng(kx + 7 * ky)
    v = rng.normal(size=ob.D) + 1j * rng.normal(size=ob.D)
    v /= np.linalg.norm(v)
    A = [1, 2, 3, 4, 6]
    M = ob.schmidt(v, A)
    n = 2 ** L
    pos, coef = ob.index(np.arange(n, dtype=np.uint64))
    P = np.zeros((n, ob.D), np.complex128)
    nz = coef != 0
    P[np.nonzero(nz)[0], pos[nz] - 1] = coef[nz] / ob.order
    psi = P.conj() @ v
    Bs = [q for q in range(1, L + 1) if q not in A]
    Md = np.zeros_like(M)
    for c in range(n):
        d = dense.digits(c, L, 2)
        Md[dense.content([d[q - 1] for q in A], 2), dense.content([d[q - 1] for q in Bs], 2)] = psi[c]
    np.testing.assert_allclose(M, Md, atol=1e-14)
    assert abs(np.linalg.norm(M) - 1) < 1e-12
