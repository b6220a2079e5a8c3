"""Custom permutation generators on the device (basis(...; symmetries),
AbelianBasis.jl:771-847, the 2D example at :754-760): 2D tori with T_x and
T_y given as site permutations (T_x in the shift-and-mask form, T_y as a
rotation of the packed word), with and without a custom spin flip.  Tables, R bits and index()
bit-exact against the oracle; the XXZ matvec to 1e-12 and the Lanczos ground
energy against the dense sector matrix."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.models import heisenberg_2d, rand_state, torus_translations

pytestmark = pytest.mark.gpu

CASES = [
    # Lx, Ly, N, kx, ky, flip (None or the flip quantum number q)
    (4, 4, 8, 0, 0, None),
    (4, 4, 8, 1, 2, None),
    (4, 4, 7, 1, 1, None),
    (4, 3, 6, 1, 2, 1),
    (6, 4, 12, 0, 0, None),
    (6, 4, 12, 2, 1, None),
    (5, 4, 10, 2, 3, 0),
]


def torus_pair(edk, Lx, Ly, N, kx, ky, flip):
    L = Lx * Ly
    tx, ty = torus_translations(Lx, Ly)
    sym = [(tx, kx), (ty, ky)]
    if flip is not None:
        sym.append((list(range(1, L + 1)), flip, [1] * L))
    return edk.basis(L=L, N=N, symmetries=sym), orc.Basis("abelian", L, N=N, symmetries=sym)


@pytest.mark.parametrize("Lx,Ly,N,kx,ky,flip", CASES)
def test_torus_tables_index_matvec(gpu, Lx, Ly, N, kx, ky, flip):
    import edkit_hip as edk
    L = Lx * Ly
    B, ob = torus_pair(edk, Lx, Ly, N, kx, ky, flip)
    I, R = B.tables()
    oI, oR = ob.tables()
    assert np.array_equal(I, oI)
    assert np.array_equal(R.view(np.uint64), oR.view(np.uint64))
    states = np.random.default_rng(L + N).integers(0, 1 << L, 1 << 14)
    coef, pos = B.index(states)
    opos, ocoef = ob.index(states.astype(np.uint64))
    assert np.array_equal(pos, opos)
    assert np.array_equal(coef.real, ocoef.real) and np.array_equal(coef.imag, ocoef.imag)
    H = heisenberg_2d(B, Lx, Ly, delta=0.8)
    x = rand_state(B.D, 3)
    y = H @ x
    yo = orc.Op(ob, H.flatten()) @ x
    assert np.abs(y - yo).max() <= 1e-12 * np.abs(yo).max()
    y32 = H @ x.astype(np.complex64)
    assert np.abs(y32 - yo).max() <= 2e-6 * np.abs(yo).max()


def test_torus_ground_state(gpu):
    """4 x 4 Heisenberg torus, (kx, ky) = (0, 0), N = 8: the device Lanczos
    energy against numpy's eigh of the oracle's sector matrix."""
    import edkit_hip as edk
    B, ob = torus_pair(edk, 4, 4, 8, 0, 0, None)
    H = heisenberg_2d(B, 4, 4)
    op = orc.Op(ob, H.flatten())
    D = ob.D
    Hs = np.stack([op @ np.eye(D, dtype=np.complex128)[:, j] for j in range(D)], axis=1)
    e0 = np.linalg.eigvalsh(0.5 * (Hs + Hs.conj().T))[0]
    E0, res, steps, _ = edk.lanczos_ground(H, rand_state(B.D, 11), m=200, tol=1e-10)
    assert abs(E0 - e0) < 1e-9 and res < 1e-8


def test_torus_double_basis_and_schmidt(gpu):
    """Custom generators through the DoubleBasis orbit walk and the Schmidt
    walker: the overlap map from a (kx, ky) torus sector into its charge
    sector, S+ between two torus sectors at zero momentum (a covariant
    operator), and the Schmidt matrix of a sector state over the first row."""
    import edkit_hip as edk
    from oracle import linearmap as olm
    Lx, Ly = 4, 3
    L = Lx * Ly
    B1, ob1 = torus_pair(edk, Lx, Ly, 6, 1, 2, None)
    Bp, obp = edk.ProjectedBasis(L=L, N=6), orc.Basis("projected", L, N=6)
    v = rand_state(Bp.D, 21)
    T = edk.DoubleBasis(B1, Bp)
    ref = olm.double_map(ob1, obp, v)
    assert np.abs(T(v) - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    Bn, obn = torus_pair(edk, Lx, Ly, 5, 0, 0, None)
    Bm, obm = torus_pair(edk, Lx, Ly, 6, 0, 0, None)
    sp = np.array([[0, 1], [0, 0]], complex)
    one = [[i] for i in range(1, L + 1)]
    H = edk.operator([sp] * L, one, edk.DoubleBasis(Bm, Bn))
    Hr = olm.double_operator(obm, obn, [sp] * L, one)
    A = edk.Array(H)
    assert A.shape == Hr.shape and np.abs(A - Hr).max() <= 1e-13 * max(1.0, np.abs(Hr).max())
    w = rand_state(B1.D, 22)
    Ainds = list(range(1, Lx + 1))
    M = edk.schmidt(w, Ainds, B1)
    Mo = ob1.schmidt(w, Ainds)
    assert np.abs(M - Mo).max() <= 1e-14
