"""CPU checks of the C ABI boundary: the library loads (no GPU needed) and
exports every function include/edkit_hip.h declares; host-side term
assembly flattens exactly like Vector{SparseMatrixCSC}."""
import ctypes
import os
import re

import numpy as np

from tests.models import Shape, j1j2, xxz

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "edkit_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(edk_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from edkit_hip import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_lib.SYMBOLS) == names


def test_version_and_error_without_gpu():
    from edkit_hip import _lib
    lib = _lib.lib()
    assert lib.edk_version() >= 1
    assert isinstance(lib.edk_last_error(), bytes)


def test_flatten_matches_csc_layout():
    H = xxz(Shape(4), 0.5)
    n, so, si, co, cp, no, rv, nz = H.flatten()
    assert n == 4  # one 4x4 term per bond, [4,1] last (Operator.jl:106)
    assert list(si[-2:]) == [4, 1]
    # XX+YY+0.5ZZ: 4 diagonal + 2 off-diagonal nonzeros per bond
    assert list(no) == [0, 6, 12, 18, 24]
    assert list(cp[:5]) == [1, 2, 4, 6, 7]
    assert np.allclose(nz[0::2][:6], [0.125, -0.125, 0.5, 0.5, -0.125, 0.125])


def test_operator_merges_equal_site_lists():
    H = j1j2(Shape(6), j2=0.3)
    assert len(H) == 12
    H2 = H + H
    assert len(H2) == 12
    assert np.allclose(H2.M[0], 2 * H.M[0])
