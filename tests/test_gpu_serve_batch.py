"""Batched matvecs of operators that share a sector and a term structure
(the server's continuous batching): the multi-coefficient block launch
(edk_matvec_multi) against the oracle column by column, the same bits per
column as alone; and the server's groups: jobs on one sector with different
couplings share their products step by step, each result the same bit for
bit alone and in the mix, and equal to the direct solves within tolerance."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.models import j1j2, rand_state, xxz

pytestmark = pytest.mark.gpu

SECTORS = [
    ("TranslationParityBasis", 16, dict(N=8, k=0, p=1), "translation_parity"),
    ("TranslationParityBasis", 18, dict(N=9, k=9, p=-1), "translation_parity"),
    ("TranslationalBasis", 14, dict(N=7, k=3), "translational"),
    ("TranslationalBasis", 20, dict(N=10, k=0), "translational"),
    ("TranslationalBasis", 12, dict(N=6, k=1, a=2), "translational"),
]


@pytest.mark.parametrize("kind,L,kw,okind", SECTORS)
@pytest.mark.parametrize("n", [1, 3, 8])
def test_matvec_multi_matches_oracle(gpu, kind, L, kw, okind, n):
    import edkit_hip as edk
    B = getattr(edk, kind)(L=L, **kw)
    ob = orc.Basis(okind, L, **kw)
    Hs = [xxz(B, 0.4 + 0.23 * c) for c in range(n)]
    xs = [rand_state(B.D, 40 + c) for c in range(n)]
    for dt, tol in ((np.complex128, 1e-12), (np.complex64, 2e-6)):
        ys = edk.mul_multi(Hs, [x.astype(dt) for x in xs])
        for c in range(n):
            yo = orc.Op(ob, Hs[c].flatten()) @ xs[c].astype(dt).astype(np.complex128)
            assert ys[c].dtype == dt
            assert np.linalg.norm(ys[c] - yo) <= tol * np.linalg.norm(yo), (dt, c)
            # the same bits as the column alone
            alone = edk.mul_multi([Hs[c]], [xs[c].astype(dt)])[0]
            assert np.array_equal(alone.view(np.float64 if dt == np.complex128 else np.float32),
                                  ys[c].view(np.float64 if dt == np.complex128 else np.float32))


def test_matvec_multi_j1j2_and_mismatch(gpu):
    import edkit_hip as edk
    B = edk.TranslationalBasis(L=18, N=9, k=4)
    ob = orc.Basis("translational", 18, N=9, k=4)
    Hs = [j1j2(B, 0.1 + 0.1 * c) for c in range(5)]
    xs = [rand_state(B.D, 60 + c) for c in range(5)]
    ys = edk.mul_multi(Hs, xs)
    for c in range(5):
        yo = orc.Op(ob, Hs[c].flatten()) @ xs[c]
        assert np.linalg.norm(ys[c] - yo) <= 1e-12 * np.linalg.norm(yo)
    # different structures (a second-neighbour family against none) are refused
    with pytest.raises(edk.EdkError):
        edk.mul_multi([Hs[0], xxz(B, 1.0)], xs[:2])


def _jobs(edk):
    """Groups: six XXZ chains (different Delta) on one L = 20 sector, four
    J1-J2 rings (different J2) on another, plus a lone job of a third."""
    B1 = edk.TranslationParityBasis(L=20, N=10, k=0, p=1)
    B2 = edk.TranslationalBasis(L=18, N=9, k=5)
    B3 = edk.TranslationalBasis(L=16, N=8, k=2)
    jobs = []
    for c in range(6):
        H = xxz(B1, 0.6 + 0.25 * c)
        if c % 2 == 0:
            jobs.append(("ground", H, dict(psi0=rand_state(B1.D, 70 + c), m=250, tol=1e-9, reorth="partial")))
        else:
            jobs.append(("evolve", H, dict(psi0=rand_state(B1.D, 70 + c), t=1.0 + 0.5 * c)))
    for c in range(4):
        H = j1j2(B2, 0.15 * (c + 1))
        jobs.append(("ground", H, dict(psi0=rand_state(B2.D, 80 + c), m=250, tol=1e-9)))
    jobs.append(("evolve", xxz(B3, 1.1), dict(psi0=rand_state(B3.D, 90), t=2.0)))
    return jobs


def _submit(srv, kind, H, kw):
    if kind == "ground":
        return srv.submit_ground(H, want_vector=True, **kw)
    return srv.submit_evolve(H, want_vector=True, **kw)


def test_server_groups_same_alone_and_in_mix(gpu):
    import edkit_hip as edk
    jobs = _jobs(edk)
    alone = []
    with edk.Server(workers=1) as srv:
        for kind, H, kw in jobs:
            alone.append(srv.wait(_submit(srv, kind, H, kw), timeout=120))
        l1, c1 = srv.batch_stats()
    assert l1 == c1  # one job at a time: blocks of one column
    with edk.Server(workers=12) as srv:
        tickets = [_submit(srv, kind, H, kw) for kind, H, kw in jobs]
        mixed = [srv.wait(t, timeout=120) for t in tickets]
        launches, columns = srv.batch_stats()
    assert columns > 1.5 * launches, (launches, columns)  # products were shared
    for (kind, H, kw), (ra, va), (rm, vm) in zip(jobs, alone, mixed):
        assert ra["status"] == 0 and rm["status"] == 0, (ra["message"], rm["message"])
        assert np.array_equal(va, vm), kind
        for f in ("E0", "resid", "steps", "matvecs", "overlap_re", "overlap_im", "norm"):
            assert ra[f] == rm[f], (kind, f)
    # against the direct solves (single-vector kernel)
    for (kind, H, kw), (ra, va) in zip(jobs, alone):
        if kind == "ground":
            E, r, _, _ = edk.lanczos_ground(H, kw["psi0"], m=kw["m"], tol=kw["tol"],
                                            reorth=kw.get("reorth", "full"))
            assert abs(E - ra["E0"]) < 1e-11 and ra["resid"] <= kw["tol"]
        else:
            ref = edk.timeevolve(H, kw["psi0"], kw["t"])
            assert np.linalg.norm(va - ref) < 1e-9


def test_server_batching_off_matches(gpu):
    """batch=False runs every product directly: the same energies within
    rounding, and no block launches."""
    import edkit_hip as edk
    jobs = [j for j in _jobs(edk) if j[0] == "ground"][:3]
    with edk.Server(workers=4, batch=False) as srv:
        res = [srv.wait(_submit(srv, k, H, kw), timeout=120)[0] for k, H, kw in jobs]
        assert srv.batch_stats() == (0, 0)
    with edk.Server(workers=4) as srv:
        res2 = [srv.wait(_submit(srv, k, H, kw), timeout=120)[0] for k, H, kw in jobs]
    for a, b in zip(res, res2):
        assert abs(a["E0"] - b["E0"]) < 1e-11
