"""GPU parity of the Schmidt decomposition of sector states (schmidt, ent_spec,
ent_S; Schmidt.jl:100-330, AbelianBasis.jl:704-719): the device Schmidt
matrix against the oracle's (ComplexF64 within 1e-14, ComplexF32 states
within FP32 rounding), the device squared Schmidt values against numpy's SVD
of the oracle matrix, the entropy conventions, and the entanglement of a
sector ground state against the same state built in the full space."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.models import rand_state, xxz

pytestmark = pytest.mark.gpu

OKIND = {"TensorBasis": "tensor", "ProjectedBasis": "projected", "TranslationalBasis": "translational",
         "TranslationParityBasis": "translation_parity", "ParityBasis": "parity", "FlipBasis": "flip",
         "ParityFlipBasis": "parity_flip", "TranslationFlipBasis": "translation_flip"}

CASES = [
    ("TensorBasis", 10, 2, {}),
    ("ProjectedBasis", 12, 2, dict(N=6)),
    ("TranslationalBasis", 14, 2, dict(N=7, k=3)),
    ("TranslationalBasis", 12, 2, dict(N=6, k=1, a=2)),
    ("TranslationParityBasis", 14, 2, dict(N=7, k=7, p=-1)),
    ("ParityBasis", 12, 2, dict(N=6, p=1)),
    ("FlipBasis", 12, 2, dict(N=6, p=-1)),
    ("ParityFlipBasis", 12, 2, dict(N=6, p=1, z=-1)),
    ("TranslationFlipBasis", 12, 2, dict(N=6, k=5, p=1)),
    ("TranslationalBasis", 8, 3, dict(N=8, k=2)),
    ("basis", 12, 2, dict(N=6, k=3, z=1)),
]


def pair(edk, kind, L, base, kw):
    if kind == "basis":
        B = edk.basis(L=L, base=base, **kw)
        ob = orc.Basis("abelian", L, base=base, N=kw.get("N"), k=kw.get("k", 0), p=kw.get("p", 0),
                       z=kw.get("z", 0), has_k="k" in kw, has_p="p" in kw, has_z="z" in kw)
        return B, ob
    return getattr(edk, kind)(L=L, base=base, **kw), orc.Basis(OKIND[kind], L, base=base, **kw)


@pytest.mark.parametrize("kind,L,base,kw", CASES)
def test_schmidt_matrix_and_spectrum_match_oracle(gpu, kind, L, base, kw):
    import edkit_hip as edk
    B, ob = pair(edk, kind, L, base, kw)
    v = rand_state(B.D, 61)
    A = [1, 3, 4, L // 2 + 1] if L > 8 else [2, 3, 7]
    Mo = ob.schmidt(v, A)
    M = edk.schmidt(v, A, B)
    assert M.shape == Mo.shape
    assert np.abs(M - Mo).max() <= 1e-14, kind
    M32 = edk.schmidt(v.astype(np.complex64), A, B)
    Mo32 = ob.schmidt(v.astype(np.complex64).astype(np.complex128), A)
    assert np.abs(M32 - Mo32).max() <= 1e-6 * np.abs(Mo32).max()
    s2 = edk.ent_spec(v, A, B) ** 2
    so = np.linalg.svd(Mo, compute_uv=False) ** 2
    assert np.abs(s2 - so).max() <= 1e-12
    for alpha in (1, 0, 2):
        assert abs(edk.ent_S(v, A, B, alpha=alpha) - edk.entropy(so, alpha=alpha)) <= 1e-11, alpha


def test_entropy_conventions():
    """entropy(s; α, cutoff) (Schmidt.jl:100-140): Shannon and support size
    stop at the first value not above the cutoff; Renyi for other α."""
    import edkit_hip as edk
    s = np.array([0.5, 0.25, 0.25, 1e-30, 0.0])
    assert abs(edk.entropy(s) - (-0.5 * np.log(0.5) - 0.5 * np.log(0.25))) < 1e-15
    assert edk.entropy(s, alpha=0) == 3
    assert abs(edk.entropy(s, alpha=2) - (-np.log(0.25 + 0.0625 + 0.0625))) < 1e-15


def test_sector_ground_state_entanglement(gpu):
    """Half-chain entanglement of the XXZ ground state (L = 12, Δ = 1): from
    the TranslationParityBasis k = 0 p = +1 sector on the device and from
    the full-space ground vector (dense, in the N = 6 sector)."""
    import edkit_hip as edk
    from tests import dense
    from tests.models import Shape
    L = 12
    B = edk.TranslationParityBasis(L=L, N=6, k=0, p=1)
    ob = orc.Basis("translation_parity", L, N=6, k=0, p=1)
    H = xxz(B)
    op = orc.Op(ob, H.flatten())
    Hs = np.stack([op @ np.eye(B.D)[:, j] for j in range(B.D)], axis=1)
    w, U = np.linalg.eigh(0.5 * (Hs + Hs.conj().T))
    v = U[:, 0]
    O = dense.full_operator(xxz(Shape(L)), L)
    I, _ = orc.Basis("projected", L, N=6).tables()
    wf, Uf = np.linalg.eigh(O[np.ix_(I - 1, I - 1)])
    assert abs(wf[0] - w[0]) < 1e-10 and wf[1] - wf[0] > 1e-6
    psi = np.zeros(2 ** L, np.complex128)
    psi[I - 1] = Uf[:, 0]
    A = list(range(1, L // 2 + 1))
    s_full = np.linalg.svd(psi.reshape(2 ** (L // 2), 2 ** (L // 2)), compute_uv=False) ** 2
    S_dev = edk.ent_S(v, A, B)
    assert abs(S_dev - edk.entropy(s_full)) < 1e-10
    assert abs(edk.ent_S(v, A, B, alpha=2) - edk.entropy(s_full, alpha=2)) < 1e-10


def test_ent_spec_at_l20(gpu):
    """A 1024 x 1024 reduced density matrix (TranslationalBasis L = 20 N = 10
    k = 0, half chain) against numpy's SVD of the oracle matrix."""
    import edkit_hip as edk
    B = edk.TranslationalBasis(L=20, N=10, k=0)
    ob = orc.Basis("translational", 20, N=10, k=0)
    v = rand_state(B.D, 62)
    A = list(range(1, 11))
    so = np.linalg.svd(ob.schmidt(v, A), compute_uv=False) ** 2
    s2 = edk.ent_spec(v, A, B) ** 2
    assert np.abs(s2 - so).max() <= 1e-12
    assert abs(edk.ent_S(v, A, B) - edk.entropy(so)) <= 1e-11


def test_ent_spec_rejects_32bit_overflow(gpu):
    """L = 32 with one site in A: the inner size 2^31 of rho_A = M M^H does not
    fit rocBLAS's 32-bit sizes, so the call is refused before any allocation
    (the dummy vector is never read)."""
    import ctypes as C
    import edkit_hip as edk
    from edkit_hip._lib import lib
    B = edk.TensorBasis(L=32)
    dummy = edk.DeviceArray((4,), np.complex128)
    A = np.array([1], np.int32)
    s2 = np.zeros(2, np.float64)
    cnt = C.c_int64(-1)
    rc = lib().edk_ent_spec(B.h, 0, dummy.ptr, 1, A.ctypes.data, s2.ctypes.data, C.byref(cnt))
    assert rc != 0 and cnt.value == -1
    assert "2^31" in lib().edk_last_error().decode()
