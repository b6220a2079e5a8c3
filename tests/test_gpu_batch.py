"""GPU parity of the batching row: block-of-vectors kernel, schedule variants,
autotuner and multi-sector graph batches, each against the oracle (ComplexF64
within 1e-12, ComplexF32 within 2e-6 relative) and against the eager
single-vector path (bitwise, where the same kernel runs)."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.models import blbq, j1j2, rand_state, xxz

pytestmark = pytest.mark.gpu

OKIND = {"ProjectedBasis": "projected", "TranslationalBasis": "translational",
         "TranslationParityBasis": "translation_parity"}

# one case per phase mode of the specialised kernel (real positive, real
# signs, complex), the unit-cell form and the projected basis
BLOCK_CASES = [
    ("TranslationParityBasis", 16, dict(N=8, k=0, p=1), xxz),
    ("TranslationParityBasis", 16, dict(N=8, k=8, p=-1), j1j2),
    ("TranslationalBasis", 16, dict(N=8, k=0), j1j2),
    ("TranslationalBasis", 16, dict(N=8, k=8), xxz),
    ("TranslationalBasis", 16, dict(N=7, k=3), j1j2),
    ("TranslationalBasis", 16, dict(N=8, k=3, a=2), xxz),
    ("ProjectedBasis", 14, dict(N=7), j1j2),
    # at most 4 targets a row: the first chunk holds the full 256 rows
    ("ProjectedBasis", 24, dict(N=2), xxz),
]


def block(D, nb, seed, dtype):
    return np.stack([rand_state(D, seed + c) for c in range(nb)], axis=1).astype(dtype)


@pytest.mark.parametrize("kind,L,kw,model", BLOCK_CASES)
@pytest.mark.parametrize("nb", [2, 4, 8])
def test_block_matvec_matches_oracle(gpu, kind, L, kw, model, nb):
    import edkit_hip as edk
    B = getattr(edk, kind)(L=L, **kw)
    ob = orc.Basis(OKIND[kind], L, **kw)
    H = model(B)
    op = orc.Op(ob, H.flatten())
    for dt, tol in ((np.complex128, 1e-12), (np.complex64, 2e-6)):
        X = block(B.D, nb, 100 * nb, dt)
        Y = H @ edk.DeviceArray.from_host(X)
        Yh = Y.to_host()
        for c in range(nb):
            yo = op @ X[:, c].astype(np.complex128)
            assert np.linalg.norm(Yh[:, c] - yo) <= tol * np.linalg.norm(yo), (kind, kw, nb, dt, c)
        # accumulate: mul!(Y, H, X) adds to Y
        Y2 = edk.DeviceArray.from_host(X)
        edk.mul_(Y2, H, edk.DeviceArray.from_host(X))
        assert np.linalg.norm(Y2.to_host() - (X + Yh)) <= tol * np.linalg.norm(X + Yh)


def test_schedule_variants_agree(gpu):
    import edkit_hip as edk
    from edkit_hip import _lib
    from edkit_hip.matvec import device_operator
    for kind, L, kw, model in BLOCK_CASES:
        B = getattr(edk, kind)(L=L, **kw)
        ob = orc.Basis(OKIND[kind], L, **kw)
        H = model(B)
        op = device_operator(H)
        yo = orc.Op(ob, H.flatten()) @ rand_state(B.D, 5)
        for dt, tol in ((np.complex128, 1e-12), (np.complex64, 2e-6)):
            x = edk.DeviceArray.from_host(rand_state(B.D, 5, dt))
            for v in (_lib.EDK_VARIANT_GENERIC, _lib.EDK_VARIANT_TILE, _lib.EDK_VARIANT_LANE2,
                      _lib.EDK_VARIANT_LANE4, _lib.EDK_VARIANT_LANE8):
                op.set_variant(dt, 1, v)
                y = (H @ x).to_host()
                assert np.linalg.norm(y - yo) <= tol * np.linalg.norm(yo), (kind, kw, dt, v)
            op.set_variant(dt, 1, _lib.EDK_VARIANT_AUTO)
    with pytest.raises(edk.EdkError):
        op.set_variant(np.complex128, 4, _lib.EDK_VARIANT_LANE2)


def test_autotune_picks_a_valid_schedule(gpu):
    import edkit_hip as edk
    from edkit_hip import _lib
    from edkit_hip.matvec import device_operator
    B = edk.TranslationParityBasis(L=20, N=10, k=0, p=1)
    ob = orc.Basis("translation_parity", 20, N=10, k=0, p=1)
    H = xxz(B)
    op = device_operator(H)
    yo = orc.Op(ob, H.flatten()) @ rand_state(B.D, 3)
    for nb in (1, 8):
        v, ms = op.autotune(np.complex128, nb, reps=3)
        assert v in _lib.VARIANT_NAMES and v != _lib.EDK_VARIANT_AUTO and ms > 0
    y = (H @ edk.DeviceArray.from_host(rand_state(B.D, 3))).to_host()
    assert np.linalg.norm(y - yo) <= 1e-12 * np.linalg.norm(yo)
    # a basis outside the specialised kernel: only the generic schedule exists
    B3 = edk.TranslationalBasis(L=8, N=8, k=0, base=3)
    v, _ = device_operator(blbq(B3)).autotune(np.complex64, 1, reps=2)
    assert v == _lib.EDK_VARIANT_GENERIC


@pytest.mark.parametrize("streams,fuse", [(1, True), (4, True), (1, False)])
def test_multi_sector_graph_batch(gpu, streams, fuse):
    """All 16 momentum sectors of a J1-J2 ring L=16 N=8 (blocks of 8, FP32)
    plus a spin-1 BLBQ sector, in one graph; every replay equals the eager
    matvecs bitwise and the oracle to FP32 rounding.  Fused, the 14 complex-
    phase sectors share two launches (at most 12 sectors a launch; k = 0,
    k = 8 and the spin-1 sector keep their own)."""
    import edkit_hip as edk
    items, eager, checks = [], [], []
    for k in range(16):
        B = edk.TranslationalBasis(L=16, N=8, k=k)
        H = j1j2(B, 0.3)
        X = edk.DeviceArray.from_host(block(B.D, 8, 7 * k, np.complex64))
        Y = edk.DeviceArray((B.D, 8), np.complex64)
        items.append((H, X, Y))
        eager.append((H @ X).to_host())
        if k in (0, 3, 8):
            checks.append((len(items) - 1, orc.Op(orc.Basis("translational", 16, N=8, k=k), H.flatten())))
    B3 = edk.TranslationalBasis(L=8, N=8, k=0, base=3)
    H3 = blbq(B3, 0.2)
    X3 = edk.DeviceArray.from_host(block(B3.D, 8, 900, np.complex64))
    Y3 = edk.DeviceArray((B3.D, 8), np.complex64)
    items.append((H3, X3, Y3))
    eager.append((H3 @ X3).to_host())
    checks.append((len(items) - 1, orc.Op(orc.Basis("translational", 8, base=3, N=8, k=0), H3.flatten())))
    batch = edk.MatvecBatch(items, streams=streams, fuse=fuse)
    assert batch.launches == (5 if fuse else 17)
    for rep in range(3):
        for _, _, Y in items:
            Y.zero_()
        batch.launch()
        gpu.synchronize()
        for (H, X, Y), ye in zip(items, eager):
            assert np.array_equal(Y.to_host(), ye), rep
        # eager work between replays
        _ = (items[0][0] @ items[0][1]).to_host()
    for i, op in checks:
        H, X, Y = items[i]
        Xh, Yh = X.to_host(), Y.to_host()
        for c in range(8):
            yo = op @ Xh[:, c].astype(np.complex128)
            assert np.linalg.norm(Yh[:, c] - yo) <= 2e-6 * np.linalg.norm(yo)
    batch.close()


def test_fused_batch_sectors_of_both_parities(gpu):
    """Fused block launches over translation-parity sectors of two lengths
    (ComplexF64, blocks of 4, accumulating): the sectors of one phase mode
    share a launch and every result equals its own sector's launch bitwise
    and the oracle to 1e-12."""
    import edkit_hip as edk
    items, eager, refs = [], [], []
    for L, k, p in ((16, 0, 1), (18, 0, 1), (16, 0, -1), (18, 9, 1), (18, 9, -1), (16, 8, 1)):
        B = edk.TranslationParityBasis(L=L, N=L // 2, k=k, p=p)
        H = xxz(B, 0.7) if L == 16 else j1j2(B, 0.4)
        X0 = block(B.D, 4, 31 * L + k + p, np.complex128)
        Y0 = block(B.D, 4, 77 * L + k + p, np.complex128)
        X = edk.DeviceArray.from_host(X0)
        Y = edk.DeviceArray.from_host(Y0)
        items.append((H, X, Y, True))
        eager.append(Y0 + (H @ X).to_host())
        refs.append((orc.Op(orc.Basis("translation_parity", L, N=L // 2, k=k, p=p), H.flatten()), X0, Y0))
    batch = edk.MatvecBatch(items, streams=1)
    # phase mode 0 (k = 0, p = +1) and mode 1 (the rest): two launches
    assert batch.launches == 2
    batch.launch()
    gpu.synchronize()
    for (H, X, Y, _), ye, (op, X0, Y0) in zip(items, eager, refs):
        Yh = Y.to_host()
        assert np.array_equal(Yh, ye)
        for c in range(4):
            yo = Y0[:, c] + op @ X0[:, c]
            assert np.linalg.norm(Yh[:, c] - yo) <= 1e-12 * np.linalg.norm(yo)
    batch.close()


def test_batch_accumulates_and_checks_arguments(gpu):
    import edkit_hip as edk
    B = edk.TranslationParityBasis(L=14, N=7, k=0, p=1)
    H = xxz(B)
    x = edk.DeviceArray.from_host(rand_state(B.D, 1))
    y0 = rand_state(B.D, 2)
    y = edk.DeviceArray.from_host(y0)
    hx = (H @ x).to_host()
    batch = edk.MatvecBatch([(H, x, y, True)], streams=2)
    batch.launch()
    batch.launch()
    gpu.synchronize()
    assert np.linalg.norm(y.to_host() - (y0 + 2 * hx)) <= 1e-13 * np.linalg.norm(y0 + 2 * hx)
    batch.close()
    with pytest.raises(ValueError):
        edk.MatvecBatch([(H, x, edk.DeviceArray((B.D,), np.complex64))])
    with pytest.raises(edk.EdkError):
        edk.MatvecBatch([(H, x, x)])


@pytest.mark.parametrize("kind,kw", [("TranslationParityBasis", dict(N=16, k=0, p=1)),
                                     ("TranslationalBasis", dict(N=16, k=0))])
def test_l32_tile_schedule_matches_generic(gpu, kind, kw):
    """L = 32 sectors whose coefficients are all positive take the run-length
    canonicaliser in the tile schedule; the generic kernel searches every
    image.  Same seeded vector, ComplexF64 and ComplexF32."""
    import edkit_hip as edk
    from edkit_hip import _lib
    from edkit_hip.matvec import device_operator
    B = getattr(edk, kind)(L=32, **kw)
    H = xxz(B)
    op = device_operator(H)
    for dt, tol in ((np.complex128, 1e-13), (np.complex64, 1e-6)):
        x = edk.DeviceArray.from_host(rand_state(B.D, 21, dt))
        op.set_variant(dt, 1, _lib.EDK_VARIANT_TILE)
        yt = (H @ x).to_host()
        op.set_variant(dt, 1, _lib.EDK_VARIANT_GENERIC)
        yg = (H @ x).to_host()
        op.set_variant(dt, 1, _lib.EDK_VARIANT_AUTO)
        assert np.linalg.norm(yt - yg) <= tol * np.linalg.norm(yg), (kind, dt)


@pytest.mark.parametrize("kind,L,kw,model", [
    ("TranslationalBasis", 32, dict(N=16, k=3), xxz),              # complex phases, L = 32
    ("TranslationParityBasis", 32, dict(N=16, k=16, p=-1), xxz),   # signs, L = 32
    ("TranslationalBasis", 28, dict(N=14, k=5), j1j2),             # complex phases, L < 32
    ("TranslationalBasis", 30, dict(N=15, k=15), xxz),             # k = pi signs, L < 32
    ("TranslationParityBasis", 26, dict(N=13, k=13, p=-1), j1j2),  # signs, L < 32
    ("TranslationParityBasis", 24, dict(N=12, k=0, p=1), j1j2),    # positive, L < 32
    ("TranslationalBasis", 21, dict(N=10, k=4), xxz),              # odd L, complex phases
    ("TranslationParityBasis", 19, dict(N=9, k=0, p=-1), j1j2),    # odd L, parity
    ("TranslationalBasis", 16, dict(N=5, k=2), j1j2),              # L = 16, low filling
    ("TranslationalBasis", 31, dict(N=11, k=7), xxz),              # L = 31
])
def test_run_length_paths_match_generic(gpu, kind, L, kw, model):
    """The run-length canonicaliser (tile kernel for one vector, block kernel
    for eight) against the generic kernel's full search, at sizes where a
    large share of targets exercise both the direct and the deferred paths."""
    import edkit_hip as edk
    from edkit_hip import _lib
    from edkit_hip.matvec import device_operator
    B = getattr(edk, kind)(L=L, **kw)
    H = model(B)
    op = device_operator(H)
    x1 = edk.DeviceArray.from_host(rand_state(B.D, 41))
    op.set_variant(np.complex128, 1, _lib.EDK_VARIANT_GENERIC)
    yg = (H @ x1).to_host()
    op.set_variant(np.complex128, 1, _lib.EDK_VARIANT_AUTO)
    yt = (H @ x1).to_host()
    assert np.linalg.norm(yt - yg) <= 1e-13 * np.linalg.norm(yg), (kind, L, kw)
    X = np.stack([rand_state(B.D, 50 + c) for c in range(8)], axis=1)
    yb = (H @ edk.DeviceArray.from_host(X)).to_host()
    op.set_variant(np.complex128, 8, _lib.EDK_VARIANT_GENERIC)
    ybg = (H @ edk.DeviceArray.from_host(X)).to_host()
    op.set_variant(np.complex128, 8, _lib.EDK_VARIANT_AUTO)
    assert np.linalg.norm(yb - ybg) <= 1e-13 * np.linalg.norm(ybg), (kind, L, kw)
