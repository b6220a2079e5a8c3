"""Pins the Bethe-ansatz ground energy (tests/bethe.py) used by the L = 32
roofline-sector Lanczos test: against dense diagonalisation of the same model
built two independent ways (the package's xxz() terms through Kronecker
embedding, and a numpy Kronecker sum), and against the known 4-site value."""
import numpy as np
import pytest

from tests.bethe import heisenberg_ring_e0
from tests.dense import full_operator
from tests.models import Shape, xxz


def _sz0_ground(Hfull, L):
    idx = [s for s in range(2 ** L) if bin(s).count("1") == L // 2]
    return np.linalg.eigvalsh(Hfull[np.ix_(idx, idx)])[0]


def test_four_site_ring_closed_form():
    assert abs(heisenberg_ring_e0(4) + 2.0) < 1e-14


@pytest.mark.parametrize("L", [4, 6, 8])
def test_bethe_matches_dense_xxz_terms(L):
    H = full_operator(xxz(Shape(L), 1.0), L)
    assert abs(_sz0_ground(H, L) - heisenberg_ring_e0(L)) < 1e-12


@pytest.mark.parametrize("L", [10, 12, 14])
def test_bethe_matches_dense_bit_construction(L):
    """The S_z = 0 block written out from bits: S^z S^z = +-1/4 per bond,
    (S^+S^- + S^-S^+)/2 swaps an antiparallel pair with amplitude 1/2."""
    idx = [s for s in range(2 ** L) if bin(s).count("1") == L // 2]
    pos = {s: n for n, s in enumerate(idx)}
    H = np.zeros((len(idx), len(idx)))
    for n, s in enumerate(idx):
        for i in range(L):
            j = (i + 1) % L
            bi, bj = (s >> i) & 1, (s >> j) & 1
            H[n, n] += 0.25 if bi == bj else -0.25
            if bi != bj:
                H[pos[s ^ (1 << i) ^ (1 << j)], n] += 0.5
    assert abs(np.linalg.eigvalsh(H)[0] - heisenberg_ring_e0(L)) < 1e-11
