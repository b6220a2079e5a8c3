"""Device-driven Lanczos steps of the Krylov evolution (full
reorthogonalisation, edk_krylov.hip) against the host-driven loop: with
EDK_KRYLOV_HOST_STEPS=1 every step synchronises and takes its decisions on
the host.  The evolved states, the diagnostics counters and the accepted
intervals must be the same bit for bit, in ComplexF64 and ComplexF32, on
sectors that extend, restart, and break down (a sector smaller than the
Krylov dimension); the states also match exp(-iHt) of the oracle's dense
sector matrix (KrylovEvolutionCache, TimeEvolution.jl:250-306, 537-611)."""
import os

import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as orc
from tests.models import j1j2, rand_state, xxz

pytestmark = pytest.mark.gpu

OKIND = {"TranslationalBasis": "translational", "TranslationParityBasis": "translation_parity"}
CASES = [
    ("TranslationalBasis", 14, dict(N=7, k=3), xxz),
    ("TranslationParityBasis", 16, dict(N=8, k=0, p=1), j1j2),
    ("TranslationalBasis", 16, dict(N=8, k=0), j1j2),
    ("TranslationParityBasis", 8, dict(N=4, k=0, p=1), xxz),  # D < m_init: breakdown
]
TIMES = [0.3, 2.0, 6.5, 15.0]


def evolve_all(edk, H, x0, dt, host):
    if host:
        os.environ["EDK_KRYLOV_HOST_STEPS"] = "1"
    else:
        os.environ.pop("EDK_KRYLOV_HOST_STEPS", None)
    try:
        c = edk.KrylovEvolutionCache(H, x0, dtype=dt)
        outs = [c.evolve(t).to_host() for t in TIMES]
        return outs, c.diagnostics
    finally:
        os.environ.pop("EDK_KRYLOV_HOST_STEPS", None)


@pytest.mark.parametrize("kind,L,kw,model", CASES)
def test_device_steps_equal_host_steps(gpu, kind, L, kw, model):
    import edkit_hip as edk
    B = getattr(edk, kind)(L=L, **kw)
    H = model(B)
    x0 = rand_state(B.D, 11 + L)
    for dt in (np.complex128, np.complex64):
        dev, ddev = evolve_all(edk, H, x0.astype(dt), dt, host=False)
        host, dhost = evolve_all(edk, H, x0.astype(dt), dt, host=True)
        for a, b in zip(dev, host):
            assert np.array_equal(a, b), (kind, kw, dt)
        assert ddev == dhost
    # against exp(-iHt) of the dense sector matrix
    op = orc.Op(orc.Basis(OKIND[kind], L, **kw), H.flatten())
    Hd = np.stack([op @ np.eye(B.D, dtype=np.complex128)[:, j] for j in range(B.D)], axis=1)
    for t, y in zip(TIMES, dev_c64(edk, H, x0)):
        ref = scipy.linalg.expm(-1j * t * Hd) @ x0
        assert np.linalg.norm(y - ref) <= 1e-8 * np.linalg.norm(ref), (kind, kw, t)


def dev_c64(edk, H, x0):
    outs, _ = evolve_all(edk, H, x0, np.complex128, host=False)
    return outs
