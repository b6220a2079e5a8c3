"""GPU parity of DoubleBasis (LinearMap.jl:88-185): the device overlap map
T(v), symmetrizer and basis_embedding against the oracle's full-space loop,
and rectangular operators between sectors (operator(..., DoubleBasis(B1, B2)))
against the oracle's column scatter, for every pair of basis kinds at L = 4
(the reference's advanced_tests.jl set) and at sizes up to L = 14, ComplexF64
and ComplexF32, one and several right-hand sides."""
import numpy as np
import pytest

from oracle import linearmap as olm
from oracle import oracle as orc
from tests.models import rand_state

pytestmark = pytest.mark.gpu

SP = np.array([[0, 1], [0, 0]], complex)

KINDS = [("TensorBasis", "tensor", {}), ("ProjectedBasis", "projected", dict(N=2)),
         ("basis", "abelian", dict(N=2, k=0, p=1)), ("ParityBasis", "parity", dict(p=1)),
         ("FlipBasis", "flip", dict(p=1, N=2)), ("ParityFlipBasis", "parity_flip", dict(p=1, z=1, N=2)),
         ("TranslationalBasis", "translational", dict(k=0, N=2)),
         ("TranslationParityBasis", "translation_parity", dict(k=0, p=1, N=2)),
         ("TranslationFlipBasis", "translation_flip", dict(k=0, p=1, N=2))]


def pair(edk, cls, okind, L, kw, base=2):
    if cls == "basis":
        B = edk.basis(L=L, base=base, **kw)
        ob = orc.Basis("abelian", L, base=base, N=kw.get("N"), k=kw.get("k", 0), p=kw.get("p", 0),
                       z=kw.get("z", 0), has_k="k" in kw, has_p="p" in kw, has_z="z" in kw)
        return B, ob
    return getattr(edk, cls)(L=L, base=base, **kw), orc.Basis(okind, L, base=base, **kw)


def test_map_every_pair_l4(gpu):
    """T(v) = symmetrizer(T) v (advanced_tests.jl:35-42) for all 81 pairs,
    device against the oracle's full-space loop."""
    import edkit_hip as edk
    L = 4
    bases = [pair(edk, c, o, L, kw) for c, o, kw in KINDS]
    rng = np.random.default_rng(40)
    for (B1, o1), (c1, _, _) in zip(bases, KINDS):
        for (B2, o2), (c2, _, _) in zip(bases, KINDS):
            T = edk.DoubleBasis(B1, B2)
            v = rng.standard_normal(B2.D) + 1j * rng.standard_normal(B2.D)
            ref = olm.double_map(o1, o2, v)
            assert np.abs(T(v) - ref).max() <= 1e-14, (c1, c2)
            assert np.abs(edk.symmetrizer(T) - olm.symmetrizer(o1, o2)).max() <= 1e-14, (c1, c2)


MAP_CASES = [
    ("TranslationalBasis", "translational", dict(N=6, k=1), "ProjectedBasis", "projected", dict(N=6)),
    ("TranslationParityBasis", "translation_parity", dict(N=6, k=0, p=-1), "TranslationalBasis", "translational",
     dict(N=6, k=0)),
    ("TensorBasis", "tensor", {}, "TranslationFlipBasis", "translation_flip", dict(N=6, k=3, p=1)),
    ("ParityFlipBasis", "parity_flip", dict(N=6, p=-1, z=1), "FlipBasis", "flip", dict(N=6, p=1)),
    ("basis", "abelian", dict(N=6, k=2, z=-1), "TensorBasis", "tensor", {}),
]


@pytest.mark.parametrize("c1,o1,kw1,c2,o2,kw2", MAP_CASES)
def test_map_l12(gpu, c1, o1, kw1, c2, o2, kw2):
    """L = 12 maps with complex momenta, a unit-cell-free abelian basis and
    the full space on either side; a block of three vectors and FP32."""
    import edkit_hip as edk
    L = 12
    B1, ob1 = pair(edk, c1, o1, L, kw1)
    B2, ob2 = pair(edk, c2, o2, L, kw2)
    T = edk.DoubleBasis(B1, B2)
    V = np.stack([rand_state(B2.D, 70 + q) for q in range(3)], axis=1)
    ref = olm.double_map(ob1, ob2, V)
    out = T(V)
    assert out.shape == (B1.D, 3)
    assert np.abs(out - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    assert np.abs(T(V[:, 1]) - ref[:, 1]).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    o32 = T(V.astype(np.complex64))
    assert np.abs(o32 - ref).max() <= 2e-6 * max(1.0, np.abs(ref).max())


def test_basis_embedding(gpu):
    import edkit_hip as edk
    B, ob = pair(edk, "TranslationParityBasis", "translation_parity", 10, dict(N=5, k=0, p=1))
    E = edk.basis_embedding(B)
    assert E.shape == (1024, B.D)
    assert np.abs(E - olm.basis_embedding(ob)).max() <= 1e-15


OP_CASES = [
    # (target, source, model): S+ between magnetisation sectors, the
    # covariant sum e^{iqj} S+_j between momenta, S+ between reflection
    # sectors, and a translation-invariant coupling from the full space into a
    # parity sector (advanced_tests.jl:47-58, DoubleBasisTest.jl:6-13)
    ("ProjectedBasis", "projected", dict(N=5), "ProjectedBasis", "projected", dict(N=4), 10, "splus"),
    ("TranslationalBasis", "translational", dict(N=6, k=3), "TranslationalBasis", "translational", dict(N=5, k=1),
     12, "splus_q2"),
    ("TranslationParityBasis", "translation_parity", dict(N=7, k=0, p=1), "TranslationParityBasis",
     "translation_parity", dict(N=6, k=0, p=1), 14, "splus"),
    ("ParityBasis", "parity", dict(p=1), "TensorBasis", "tensor", {}, 8, "herm2"),
    ("TranslationFlipBasis", "translation_flip", dict(N=5, k=2, p=-1), "TranslationFlipBasis", "translation_flip",
     dict(N=5, k=2, p=-1), 10, "xxz"),
]


def model(name, L):
    one = [[i] for i in range(1, L + 1)]
    if name == "splus":
        return [SP] * L, one
    if name == "splus_q2":
        return [np.exp(2j * np.pi * 2 * i / L) * SP for i in range(1, L + 1)], one
    if name == "herm2":
        rng = np.random.default_rng(9)
        m = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
        m = m + m.conj().T
        return [m] * L, [[i, i % L + 1] for i in range(1, L + 1)]
    sx = np.array([[0, 1], [1, 0]]) / 2
    sy = np.array([[0, -1j], [1j, 0]]) / 2
    sz = np.array([[1, 0], [0, -1]]) / 2
    m = np.kron(sx, sx) + np.kron(sy, sy) + 0.7 * np.kron(sz, sz)
    return [m] * L, [[i, i % L + 1] for i in range(1, L + 1)]


@pytest.mark.parametrize("c1,o1,kw1,c2,o2,kw2,L,name", OP_CASES)
def test_rectangular_operator(gpu, c1, o1, kw1, c2, o2, kw2, L, name):
    import edkit_hip as edk
    B1, ob1 = pair(edk, c1, o1, L, kw1)
    B2, ob2 = pair(edk, c2, o2, L, kw2)
    mats, inds = model(name, L)
    H = edk.operator(mats, inds, edk.DoubleBasis(B1, B2))
    assert H.shape == (B1.D, B2.D)
    ref = olm.double_operator(ob1, ob2, mats, inds)
    A = edk.Array(H)
    assert A.shape == ref.shape
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(A - ref).max() <= 1e-13 * scale
    X = np.stack([rand_state(B2.D, 80 + q) for q in range(3)], axis=1)
    assert np.abs(H @ X - ref @ X).max() <= 1e-12 * scale * np.sqrt(B2.D)
    y = rand_state(B1.D, 90)
    y0 = y.copy()
    edk.mul_(y, H, X[:, 0])
    assert np.abs(y - (y0 + ref @ X[:, 0])).max() <= 1e-12 * scale * np.sqrt(B2.D)
    y32 = H @ X[:, 0].astype(np.complex64)
    assert np.abs(y32 - ref @ X[:, 0]).max() <= 2e-6 * scale * np.sqrt(B2.D)


def test_trans_inv_operator_shapes(gpu):
    """DoubleBasisTest.jl:6-13: an empty source sector gives a 5 x 0 matrix;
    advanced_tests.jl:47-49: full space into a parity sector."""
    import edkit_hip as edk
    b1 = edk.TranslationalBasis(N=2, k=0, L=10)
    b2 = edk.TranslationalBasis(N=0, k=1, L=10)
    m = np.array([[0.3, 0.1 - 0.2j], [0.1 + 0.2j, -0.5]])
    assert edk.Array(edk.trans_inv_operator(m, 1, edk.DoubleBasis(b1, b2))).shape == (5, 0)
    bpar, bfull = edk.ParityBasis(L=4, p=1), edk.TensorBasis(L=4)
    H = edk.Array(edk.trans_inv_operator(m, 1, edk.DoubleBasis(bpar, bfull)))
    assert H.shape == (bpar.D, 16)
    ob = orc.Basis("parity", 4, p=1)
    ref = olm.double_operator(ob, orc.Basis("tensor", 4), [m] * 4, [[i] for i in range(1, 5)])
    assert np.abs(H - ref).max() <= 1e-14
    with pytest.raises(ValueError, match="same length"):
        edk.DoubleBasis(edk.TensorBasis(L=3), edk.TensorBasis(L=4))
    with pytest.raises(ValueError, match="same local base"):
        edk.DoubleBasis(edk.TensorBasis(L=4), edk.TensorBasis(L=4, base=3))


def test_spin1_transition_weights(gpu):
    """DoubleBasisTest.jl:15-46 on the device: rectangular spin-1 operators
    between every pair of neighbouring digit-sum sectors (ProjectedBasis,
    base 3, L = 6), summed transition weights equal to the full-space sum."""
    import edkit_hip as edk
    L = 6
    rng = np.random.default_rng(15)
    p3, p2, p1, z0 = [1], [2, 4, 10], [3, 5, 7, 11, 13, 19], [6, 8, 12, 14, 16, 20, 22]
    m1, m2, m3 = [9, 15, 17, 21, 23, 25], [18, 24, 26], [27]
    Sp = np.zeros((27, 27), complex)
    for a, b in ((m2, m3), (m1, m2), (z0, m1), (p1, z0), (p2, p1), (p3, p2)):
        Sp[np.ix_(np.array(a) - 1, np.array(b) - 1)] = rng.random((len(a), len(b))) + 1j * rng.random((len(a), len(b)))
    XY = edk.spin((1, "+-"), (1, "-+"), (1, "z1"), (1, "1z"), D=3)
    full = edk.TensorBasis(L=L, base=3)
    Hf = edk.Array(edk.trans_inv_operator(XY, [1, 2], full))
    Of = edk.Array(edk.trans_inv_operator(Sp, [1, 2, 3], full))
    _, v = np.linalg.eigh(0.5 * (Hf + Hf.conj().T))
    B = np.sum(np.abs(v.conj().T @ Of @ v) ** 2)
    A = 0.0
    # digit sum n is N = 2L - n (N counts base - 1 - digit); the predicate
    # of the reference test is applied too (host-filtered, uploaded tables)
    for n in range(1, 2 * L + 1):
        b1 = edk.ProjectedBasis(L=L, base=3, N=2 * L - (n - 1), f=lambda x, n=n: sum(x) == n - 1)
        b2 = edk.ProjectedBasis(L=L, base=3, N=2 * L - n, f=lambda x, n=n: sum(x) == n)
        e1, v1 = np.linalg.eigh(edk.Array(edk.trans_inv_operator(XY, [1, 2], b1)))
        e2, v2 = np.linalg.eigh(edk.Array(edk.trans_inv_operator(XY, [1, 2], b2)))
        op = edk.Array(edk.trans_inv_operator(Sp, [1, 2, 3], edk.DoubleBasis(b1, b2)))
        A += np.sum(np.abs(v1.conj().T @ op @ v2) ** 2)
    assert abs(A - B) <= 1e-7 * max(1.0, B)
