"""The basis-level utilities content(b, i), norm(b, i), change!(b, i) and
index(b, dgt) (AbstractBasis.jl:95-107, 160-176, 240-246, and each basis'
index method) against the oracle's tables and index() on the same bases:
labels and norms bit-exact, change! -> index round trips, and index() of
seeded random digit rows equal to the oracle's (coefficient and position)."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.test_gpu_parity import make_pair

pytestmark = pytest.mark.gpu

CASES = [
    ("TranslationParityBasis", 12, 2, dict(N=6, k=0, p=1)),
    ("TranslationalBasis", 10, 2, dict(N=5, k=3)),
    ("TranslationalBasis", 6, 3, dict(N=6, k=2)),
    ("ProjectedBasis", 10, 2, dict(N=5)),
    ("basis", 10, 2, dict(N=5, k=0, p=-1, z=1)),
    ("FlipBasis", 10, 2, dict(N=5, p=-1)),
]


@pytest.mark.parametrize("kind,L,base,kw", CASES)
def test_content_norm_change_index(gpu, kind, L, base, kw):
    import edkit_hip as edk
    B, ob = make_pair(edk, kind, L, base, **kw)
    oI, oR = ob.tables()
    reduced = kind != "ProjectedBasis"
    for i in range(1, B.D + 1):
        assert edk.content(B, i) == oI[i - 1]
        nrm = edk.norm(B, i)
        assert nrm == (oR[i - 1] if reduced else 1.0)
        d, nrm2 = edk.change_(B, i)
        assert nrm2 == nrm
        assert B.content(d) == oI[i - 1] - 1
        coef, pos = edk.index(B, d)
        assert pos == i and coef == nrm
    rng = np.random.default_rng(L + base)
    for _ in range(200):
        d = rng.integers(0, base, L)
        coef, pos = edk.index(B, d)
        opos, ocoef = ob.index(np.array([B.content(d)], np.uint64))
        assert pos == opos[0]
        assert complex(coef) == complex(ocoef[0])
    with pytest.raises(IndexError):
        edk.content(B, 0)
    with pytest.raises(IndexError):
        edk.norm(B, B.D + 1)
    with pytest.raises(ValueError):
        edk.index(B, np.full(L, base))


def test_tensor_basis_utilities(gpu):
    import edkit_hip as edk
    B = edk.TensorBasis(L=6, base=3)
    for i in (1, 2, 100, 3 ** 6):
        assert edk.content(B, i) == i and edk.norm(B, i) == 1.0
        d, _ = edk.change_(B, i)
        assert edk.index(B, d) == (1.0, i)
