"""The Abelian two-site kernel (edk_perm.hip): Abelian bases of spin-1/2
sites with site-permutation generators (basis(...; symmetries),
AbelianBasis.jl:771-847) and operators made of two-site terms, on the tile
schedule, against the oracle's sector product (Operator.jl:417-465 over the
canonical forms of AbelianBasis.jl:428-445) and against the generic kernel.

Cases cover the pruned walk (generator 0 maps the first row onto itself)
and the full walk (generator 0 first in the list does not, or the pruning
table is too large for LDS), real and complex phases, a third (spin-flip)
generator, bonds listed in either orientation, four families (next-nearest
neighbours with their own matrix) and repeated bonds, keyword bases (one
translation generator, a reflection, a spin flip) and a ring translation
given as a permutation; an operator with one-site terms, and a basis
without a fixed charge, keep the generic kernel."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.models import heisenberg_2d, rand_state, torus_sites, torus_translations

pytestmark = pytest.mark.gpu


def torus(edk, Lx, Ly, N, kx, ky, flip=None, ty_first=False):
    L = Lx * Ly
    tx, ty = torus_translations(Lx, Ly)
    sym = [(ty, ky), (tx, kx)] if ty_first else [(tx, kx), (ty, ky)]
    if flip is not None:
        sym.append((list(range(1, L + 1)), flip, [1] * L))
    return edk.basis(L=L, N=N, symmetries=sym), orc.Basis("abelian", L, N=N, symmetries=sym)


def heis(delta=1.0):
    from edkit_hip.spin import spin
    return spin((1.0, "xx"), (1.0, "yy"), (delta, "zz"))


def j1j2_2d(B, Lx, Ly, j2=0.4, reversed_bonds=False):
    """Nearest neighbours (x and y) plus both diagonals with J2; with
    reversed_bonds every nearest-neighbour bond lists its neighbour first."""
    from edkit_hip.operator import operator
    h1, h2 = heis(0.9), j2 * heis(0.9)
    mats, bonds = [], []
    for (x, y) in torus_sites(Lx, Ly):
        s = 1 + x + Lx * y
        nx = 1 + (x + 1) % Lx + Lx * y
        ny = 1 + x + Lx * ((y + 1) % Ly)
        for nb in (nx, ny):
            mats.append(h1)
            bonds.append([nb, s] if reversed_bonds else [s, nb])
        for nb in (1 + (x + 1) % Lx + Lx * ((y + 1) % Ly), 1 + (x - 1) % Lx + Lx * ((y + 1) % Ly)):
            mats.append(h2)
            bonds.append([s, nb])
    return operator(mats, bonds, B)


def check_paths(edk, B, ob, H, seed):
    """Tile (the two-site kernel) and generic schedules against the oracle,
    ComplexF64 to 1e-12 and ComplexF32 to 2e-6, and the accumulating form."""
    from edkit_hip import _lib
    from edkit_hip.matvec import device_operator
    op = device_operator(H)
    x = rand_state(B.D, seed)
    yo = orc.Op(ob, H.flatten()) @ x
    for dt, tol in ((np.complex128, 1e-12), (np.complex64, 2e-6)):
        for v in (_lib.EDK_VARIANT_TILE, _lib.EDK_VARIANT_GENERIC):
            op.set_variant(dt, 1, v)
            y = (H @ edk.DeviceArray.from_host(x.astype(dt))).to_host()
            assert np.linalg.norm(y - yo) <= tol * np.linalg.norm(yo), (dt, v)
        op.set_variant(dt, 1, _lib.EDK_VARIANT_AUTO)
    y0 = rand_state(B.D, seed + 1)
    Y = edk.DeviceArray.from_host(y0)
    edk.mul_(Y, H, edk.DeviceArray.from_host(x))
    assert np.linalg.norm(Y.to_host() - (y0 + yo)) <= 1e-12 * np.linalg.norm(y0 + yo)


TORI = [
    # Lx, Ly, N, kx, ky, flip, ty_first
    (4, 4, 8, 0, 0, None, False),
    (4, 4, 7, 1, 2, None, False),
    (4, 3, 6, 1, 2, 1, False),
    (5, 4, 10, 2, 3, 0, False),
    (6, 4, 12, 2, 1, None, False),
    (6, 5, 7, 0, 0, None, False),
    (6, 5, 7, 1, 2, None, False),
    (4, 4, 8, 1, 0, None, True),
    (6, 4, 11, 0, 2, None, True),
    (10, 3, 6, 0, 1, None, False),
    (4, 2, 4, 1, 1, None, False),
]


@pytest.mark.parametrize("Lx,Ly,N,kx,ky,flip,ty_first", TORI)
def test_heisenberg_torus(gpu, Lx, Ly, N, kx, ky, flip, ty_first):
    import edkit_hip as edk
    B, ob = torus(edk, Lx, Ly, N, kx, ky, flip, ty_first)
    check_paths(edk, B, ob, heisenberg_2d(B, Lx, Ly, delta=0.8), 100 + Lx * Ly + N)


@pytest.mark.parametrize("rev", [False, True])
@pytest.mark.parametrize("Lx,Ly,N,kx,ky", [(4, 4, 8, 0, 0), (6, 4, 12, 1, 2), (5, 4, 9, 2, 0)])
def test_j1j2_torus_four_families(gpu, Lx, Ly, N, kx, ky, rev):
    import edkit_hip as edk
    B, ob = torus(edk, Lx, Ly, N, kx, ky)
    check_paths(edk, B, ob, j1j2_2d(B, Lx, Ly, reversed_bonds=rev), 300 + Lx * Ly + N)


def test_one_site_terms_keep_the_generic_kernel(gpu):
    import edkit_hip as edk
    from edkit_hip import _lib
    from edkit_hip.matvec import device_operator
    from edkit_hip.operator import operator
    Lx, Ly = 4, 4
    B, ob = torus(edk, Lx, Ly, 8, 1, 1)
    H = heisenberg_2d(B, Lx, Ly)
    op = device_operator(H)
    op.set_variant(np.complex128, 1, _lib.EDK_VARIANT_TILE)  # the two-site kernel applies
    mats = [heis()] * (2 * Lx * Ly) + [np.diag([0.5, -0.5]).astype(complex)] * (Lx * Ly)
    bonds = [[1 + x + Lx * y, 1 + (x + 1) % Lx + Lx * y] for (x, y) in torus_sites(Lx, Ly)]
    bonds += [[1 + x + Lx * y, 1 + x + Lx * ((y + 1) % Ly)] for (x, y) in torus_sites(Lx, Ly)]
    bonds += [[i] for i in range(1, Lx * Ly + 1)]
    H2 = operator(mats, bonds, B)
    with pytest.raises(edk.EdkError):
        device_operator(H2).set_variant(np.complex128, 1, _lib.EDK_VARIANT_TILE)
    x = rand_state(B.D, 7)
    yo = orc.Op(ob, H2.flatten()) @ x
    y = H2 @ x
    assert np.linalg.norm(y - yo) <= 1e-12 * np.linalg.norm(yo)


def keyword_pair(edk, L, **kw):
    """basis(...) with k/p/z keywords (translation, reflection and spin-flip
    generators of the Abelian walk) and its oracle twin."""
    B = edk.basis(L=L, **kw)
    ob = orc.Basis("abelian", L, N=kw.get("N"), k=kw.get("k") or 0, p=kw.get("p") or 0, z=kw.get("z") or 0,
                   has_k="k" in kw, has_p="p" in kw, has_z="z" in kw)
    return B, ob


@pytest.mark.parametrize("kw", [dict(N=7, k=2), dict(N=7, k=0, p=1, z=1), dict(N=7, k=7, p=-1, z=-1),
                                dict(N=6, k=6, p=1)])
def test_keyword_abelian_ring(gpu, kw):
    """One generator (a translation: unpruned walk), a reflection (bit
    reversal) and a spin flip (inversion) through the two-site kernel, on the
    XXZ and J1-J2 rings.  k and p together only at k = 0 or pi, the sectors
    where translation and reflection commute (the reference's notes on
    basis(...), AbelianBasis.jl:749-750)."""
    import edkit_hip as edk
    from tests.models import j1j2, xxz
    L = 14 if kw["N"] == 7 else 12
    B, ob = keyword_pair(edk, L, **kw)
    check_paths(edk, B, ob, xxz(B, 0.7), 500 + L)
    check_paths(edk, B, ob, j1j2(B, 0.35), 600 + L)


def test_custom_translation_single_generator(gpu):
    """A ring translation given as a site permutation (the only generator)."""
    import edkit_hip as edk
    L, N, k = 14, 7, 3
    t = [(i + 1) % L + 1 for i in range(L)]
    sym = [(t, k)]
    B = edk.basis(L=L, N=N, symmetries=sym)
    ob = orc.Basis("abelian", L, N=N, symmetries=sym)
    from edkit_hip.operator import operator
    H = operator([heis(0.6)] * L, [[i + 1, (i + 1) % L + 1] for i in range(L)], B)
    check_paths(edk, B, ob, H, 700)


def test_without_fixed_charge_keeps_the_generic_kernel(gpu):
    """Without N the representatives' popcounts differ: no rank directory,
    so the generic kernel runs (and still matches the oracle)."""
    import edkit_hip as edk
    from edkit_hip import _lib
    from edkit_hip.matvec import device_operator
    Lx, Ly = 4, 3
    tx, ty = torus_translations(Lx, Ly)
    sym = [(tx, 1), (ty, 0)]
    B = edk.basis(L=Lx * Ly, symmetries=sym)
    ob = orc.Basis("abelian", Lx * Ly, symmetries=sym)
    H = heisenberg_2d(B, Lx, Ly, delta=0.8)
    with pytest.raises(edk.EdkError):
        device_operator(H).set_variant(np.complex128, 1, _lib.EDK_VARIANT_TILE)
    x = rand_state(B.D, 9)
    yo = orc.Op(ob, H.flatten()) @ x
    y = H @ x
    assert np.linalg.norm(y - yo) <= 1e-12 * np.linalg.norm(yo)
