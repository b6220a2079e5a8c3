"""GPU parity of the explicit sector matrix (sparse(opt), Array(opt), addto!,
the sparse!(opt) cache, Operator.jl:171-285): the device-assembled matrix against
the oracle's sector matrix column by column (the oracle's scatter mul on unit
vectors), within 1e-12; products with several right-hand sides through the
cached matrix against the matrix-free kernels."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.models import blbq, j1j2, rand_state, xxz

pytestmark = pytest.mark.gpu

CASES = [
    ("TensorBasis", 10, 2, {}, xxz),
    ("ProjectedBasis", 12, 2, dict(N=6), j1j2),
    ("TranslationalBasis", 12, 2, dict(N=6, k=5), j1j2),
    ("TranslationParityBasis", 12, 2, dict(N=6, k=0, p=-1), xxz),
    ("TranslationalBasis", 8, 3, dict(N=8, k=2), blbq),
    ("basis", 12, 2, dict(N=6, k=2, z=-1), xxz),
]
OKIND = {"TensorBasis": "tensor", "ProjectedBasis": "projected", "TranslationalBasis": "translational",
         "TranslationParityBasis": "translation_parity"}


def pair(edk, kind, L, base, kw):
    if kind == "basis":
        B = edk.basis(L=L, base=base, **kw)
        ob = orc.Basis("abelian", L, base=base, N=kw.get("N"), k=kw.get("k", 0), p=kw.get("p", 0),
                       z=kw.get("z", 0), has_k="k" in kw, has_p="p" in kw, has_z="z" in kw)
        return B, ob
    return getattr(edk, kind)(L=L, base=base, **kw), orc.Basis(OKIND[kind], L, base=base, **kw)


@pytest.mark.parametrize("kind,L,base,kw,model", CASES)
def test_assembled_matrix_matches_oracle(gpu, kind, L, base, kw, model):
    import edkit_hip as edk
    B, ob = pair(edk, kind, L, base, kw)
    H = model(B)
    S = edk.sparse(H)
    assert S.shape == (B.D, B.D)
    # rows sorted by column, no duplicate columns, no stored zeros
    for i in range(B.D):
        cols = S.indices[S.indptr[i]:S.indptr[i + 1]]
        assert np.all(np.diff(cols) > 0)
    assert np.all(S.data != 0)
    op = orc.Op(ob, H.flatten())
    Hd = edk.Array(H)
    for j in range(B.D):
        e = np.zeros(B.D, np.complex128)
        e[j] = 1.0
        col = op @ e
        assert np.linalg.norm(Hd[:, j] - col) <= 1e-12 * max(1.0, np.linalg.norm(col)), (kind, kw, j)
    edk.clear_sparse_cache_(H)


def test_cached_matrix_serves_block_products(gpu):
    import edkit_hip as edk
    B = edk.TranslationParityBasis(L=18, N=9, k=0, p=1)
    H = j1j2(B, 0.4)
    X = np.stack([rand_state(B.D, 30 + c) for c in range(4)], axis=1)
    Xd = edk.DeviceArray.from_host(X)
    y_free = (H @ Xd).to_host()
    S = edk.sparse_(H)
    # the cached matrix now serves the block product; a single vector stays matrix-free
    y_csr = (H @ Xd).to_host()
    assert np.linalg.norm(y_csr - y_free) <= 1e-12 * np.linalg.norm(y_free)
    assert np.linalg.norm(S @ X - y_free) <= 1e-12 * np.linalg.norm(y_free)
    Y = edk.DeviceArray.from_host(X)
    edk.mul_(Y, H, Xd)  # accumulate
    assert np.linalg.norm(Y.to_host() - (X + y_free)) <= 1e-12 * np.linalg.norm(X + y_free)
    for dt in (np.complex64,):
        Xs = edk.DeviceArray.from_host(X.astype(dt))
        ys = (H @ Xs).to_host()
        assert np.linalg.norm(ys - y_free) <= 2e-6 * np.linalg.norm(y_free)
    edk.clear_sparse_cache_()
    assert np.linalg.norm((H @ Xd).to_host() - y_free) <= 1e-12 * np.linalg.norm(y_free)
    # Hermitian sector matrix
    assert abs(S - S.conj().T).max() <= 1e-12


def test_addto_accumulates_in_place(gpu):
    """addto!(M, opt) adds the sector matrix into M and returns M: a dense array
    in place, a CSR and a CSC matrix through their storage; shape and element
    type mismatches raise."""
    import scipy.sparse as sp
    import edkit_hip as edk
    B, ob = pair(edk, "TranslationalBasis", 12, 2, dict(N=6, k=5))
    H = j1j2(B, 0.3)
    op = orc.Op(ob, H.flatten())
    ref = np.stack([op @ np.eye(B.D, dtype=np.complex128)[:, j] for j in range(B.D)], axis=1)
    rng = np.random.default_rng(8)
    M0 = rng.standard_normal((B.D, B.D)) + 1j * rng.standard_normal((B.D, B.D))
    M = M0.copy()
    assert edk.addto_(M, H) is M
    assert np.abs(M - (M0 + ref)).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    for fmt in ("csr", "csc"):
        S0 = sp.random(B.D, B.D, density=0.05, random_state=3, format=fmt) * (1 + 0.5j)
        S = S0.copy()
        assert edk.addto_(S, H) is S and S.format == fmt
        assert np.abs(S.toarray() - (S0.toarray() + ref)).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    with pytest.raises(ValueError):
        edk.addto_(np.zeros((B.D + 1, B.D), complex), H)
    with pytest.raises(TypeError):
        edk.addto_(sp.csr_matrix((B.D, B.D), dtype=np.float64), H)
