"""The generic gather kernel's two schedules (one lane per row, and the tile
form with the targets compacted into LDS) against the oracle's scatter mul
(Operator.jl:399-465) on the same seeded inputs, for every basis kind that
reaches them: ComplexF64 to 1e-12, the accumulating mul! form, ComplexF32 to
2e-6, the two schedules against each other, and blocks of vectors.
EDK_GENERIC_SCHEDULE picks
the schedule when an operator is first used, so each schedule gets an
operator of its own."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.models import blbq, heisenberg_2d, j1j2, rand_state, torus_translations, xxz
from tests.test_gpu_parity import make_pair

pytestmark = pytest.mark.gpu

CASES = [
    ("TranslationParityBasis", 14, 2, dict(N=7, k=0, p=1), xxz),
    ("TranslationalBasis", 14, 2, dict(N=7, k=3), j1j2),
    ("TranslationalBasis", 10, 3, dict(N=10, k=1), blbq),
    ("ParityBasis", 12, 2, dict(N=6, p=-1), xxz),
    ("basis", 12, 2, dict(N=6, k=0, p=-1), j1j2),
    ("ProjectedBasis", 14, 2, dict(N=7), j1j2),
    ("FlipBasis", 14, 2, dict(N=7, p=1), xxz),
    ("TranslationFlipBasis", 8, 3, dict(N=8, k=3, p=1), blbq),
    ("torus", 16, 2, dict(N=8, k=(1, 2)), None),
]


def _pair(edk, kind, L, base, kw):
    if kind == "torus":
        tx, ty = torus_translations(4, 4)
        sym = [(tx, kw["k"][0]), (ty, kw["k"][1])]
        return (edk.basis(L=L, N=kw["N"], symmetries=sym),
                orc.Basis("abelian", L, N=kw["N"], symmetries=sym))
    return make_pair(edk, kind, L, base, **kw)


@pytest.mark.parametrize("kind,L,base,kw,model", CASES)
def test_generic_schedules_match_oracle(gpu, monkeypatch, kind, L, base, kw, model):
    import edkit_hip as edk
    from edkit_hip import _lib
    from edkit_hip.matvec import device_operator
    B, ob = _pair(edk, kind, L, base, kw)
    build = (lambda: heisenberg_2d(B, 4, 4, delta=0.8)) if kind == "torus" else (lambda: model(B))
    x = rand_state(B.D, 31)
    yo = orc.Op(ob, build().flatten()) @ x
    yo32 = orc.Op(ob, build().flatten()) @ x.astype(np.complex64).astype(np.complex128)
    out = {}
    for sched in ("row", "tile"):
        monkeypatch.setenv("EDK_GENERIC_SCHEDULE", sched)
        H = build()
        for dt in (np.complex128, np.complex64):
            device_operator(H).set_variant(dt, 1, _lib.EDK_VARIANT_GENERIC)
        y = H @ x
        assert np.linalg.norm(y - yo) <= 1e-12 * np.linalg.norm(yo), sched
        y2 = y.copy()
        edk.mul_(y2, H, x)
        assert np.linalg.norm(y2 - 2 * yo) <= 1e-12 * np.linalg.norm(yo), sched
        y32 = H @ x.astype(np.complex64)
        assert np.linalg.norm(y32 - yo32) <= 2e-6 * np.linalg.norm(yo32), sched
        out[sched] = y
    assert np.abs(out["row"] - out["tile"]).max() <= 1e-13 * np.abs(yo).max()


@pytest.mark.parametrize("kind,L,base,kw,model", CASES)
def test_generic_schedules_blocks(gpu, monkeypatch, kind, L, base, kw, model):
    """Blocks of 3 (a 2-column group and a single column) and 8 row-interleaved
    vectors through both schedules: every column against the oracle's mul on
    that column, ComplexF64 and ComplexF32, and the accumulating form."""
    import edkit_hip as edk
    B, ob = _pair(edk, kind, L, base, kw)
    build = (lambda: heisenberg_2d(B, 4, 4, delta=0.8)) if kind == "torus" else (lambda: model(B))
    op = orc.Op(ob, build().flatten())
    for nbk in (3, 8):
        X = np.stack([rand_state(B.D, 40 + c) for c in range(nbk)], axis=1)
        ref = np.stack([op @ X[:, c] for c in range(nbk)], axis=1)
        ref32 = np.stack([op @ X[:, c].astype(np.complex64).astype(np.complex128) for c in range(nbk)], axis=1)
        scale = np.linalg.norm(ref)
        for sched in ("row", "tile"):
            monkeypatch.setenv("EDK_GENERIC_SCHEDULE", sched)
            H = build()
            Y = (H @ edk.DeviceArray.from_host(X)).to_host()
            assert np.linalg.norm(Y - ref) <= 1e-12 * scale, (sched, nbk)
            Yacc = edk.DeviceArray.from_host(X)
            edk.mul_(Yacc, H, edk.DeviceArray.from_host(X))
            assert np.linalg.norm(Yacc.to_host() - (X + ref)) <= 1e-12 * np.linalg.norm(X + ref), (sched, nbk)
            Y32 = (H @ edk.DeviceArray.from_host(X.astype(np.complex64))).to_host()
            assert np.linalg.norm(Y32 - ref32) <= 2e-6 * np.linalg.norm(ref32), (sched, nbk)


@pytest.mark.parametrize("L,base,kw,model", [(12, 2, dict(N=6, k=3), j1j2), (8, 3, dict(N=8, k=2), blbq)])
def test_uploaded_tables_through_tile_kernel(gpu, L, base, kw, model):
    """A predicate that keeps every state makes the basis upload its tables, so
    the norms R_j are gathered from the table instead of recomputed from the
    orbit (and the specialised kernels do not apply).  One vector and a block of
    4 through the tile kernel against the oracle on the unfiltered sector."""
    import edkit_hip as edk
    B = edk.TranslationalBasis(L=L, base=base, f=lambda d: True, **kw)
    ob = orc.Basis("translational", L, base=base, **kw)
    oI, oR = ob.tables()
    I, R = B.tables()
    assert np.array_equal(I, oI) and np.array_equal(R, oR)
    H = model(B)
    op = orc.Op(ob, H.flatten())
    x = rand_state(B.D, 51)
    yo = op @ x
    assert np.linalg.norm(H @ x - yo) <= 1e-12 * np.linalg.norm(yo)
    X = np.stack([rand_state(B.D, 52 + c) for c in range(4)], axis=1)
    ref = np.stack([op @ X[:, c] for c in range(4)], axis=1)
    Y = (H @ edk.DeviceArray.from_host(X)).to_host()
    assert np.linalg.norm(Y - ref) <= 1e-12 * np.linalg.norm(ref)
