"""Bases: sector enumeration on the device behind the reference constructors.

TensorBasis           AbstractBasis.jl:217-247
ProjectedBasis        ProjectedBasis.jl:244-260
TranslationalBasis    TranslationalBasis.jl:235-266
TranslationParityBasis TranslationalParityBasis.jl:141-175
ParityBasis           ParityBasis.jl:76-97
FlipBasis             FlipBasis.jl:72-95
ParityFlipBasis       ParityFlipBasis.jl:87-111
TranslationFlipBasis  TranslationalFlipBasis.jl:122-158
basis(...)            AbelianBasis.jl:771-847
A predicate `f` is evaluated on the host over the device-built table (the
reference evaluates it on the unshifted candidate, i.e. the representative,
before its orbit judge, so filtering the representatives is the same set).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import BasisSpec, check, lib
from .device import default_context


class AbstractBasis:
    complex_phases = False

    def __init__(self, kind, L, base=2, a=1, N=None, k=0, p=1, z=0, has_k=False,
                 has_p=False, has_z=False, symmetries=None, f=None, ctx=None):
        self.ctx = ctx or default_context()
        self.kind = int(kind)
        self.L, self.base, self.a = int(L), int(base), int(a)
        self.N, self.k, self.p, self.z = N, k, p, z
        spec = BasisSpec(kind, self.L, self.base, self.a, -1 if N is None else int(N),
                         int(k or 0), int(p or 0), int(z or 0), int(has_k), int(has_p), int(has_z),
                         0, None, None, None)
        keep = []
        if symmetries:
            perms = np.ascontiguousarray(np.concatenate(
                [np.asarray(s[0], np.int32) for s in symmetries]))
            ks = np.ascontiguousarray(np.asarray([s[1] for s in symmetries], np.int32))
            invs = np.ascontiguousarray(np.concatenate([
                np.asarray(s[2] if len(s) > 2 else np.zeros(self.L), np.uint8) for s in symmetries]))
            spec.ncustom = len(symmetries)
            spec.perms, spec.ks, spec.invs = perms.ctypes.data, ks.ctypes.data, invs.ctypes.data
            keep = [perms, ks, invs]
        self._spec = spec
        h = C.c_void_p()
        check(lib().edk_basis_build(self.ctx.h, C.byref(spec), C.byref(h)))
        self.h = h
        if f is not None:
            self._apply_predicate(f)
        del keep

    def _apply_predicate(self, f):
        I, R = self.tables()
        keep = np.fromiter((bool(f(self.digits(c - 1))) for c in I), bool, len(I))
        I, R = np.ascontiguousarray(I[keep]), np.ascontiguousarray(R[keep])
        h = C.c_void_p()
        spec = self._spec
        spec.ncustom = 0
        check(lib().edk_basis_upload(self.ctx.h, C.byref(spec), I.ctypes.data, R.ctypes.data,
                                     len(I), C.byref(h)))
        lib().edk_basis_destroy(self.h)
        self.h = h

    def digits(self, content0):
        """change!(dgt, content+1) (AbstractBasis.jl:348-360): big-endian digits."""
        d = np.zeros(self.L, np.int64)
        c = int(content0)
        for i in range(self.L - 1, -1, -1):
            c, d[i] = divmod(c, self.base)
        return d

    @property
    def D(self):
        return lib().edk_basis_dim(self.h)

    def size(self, dim=None):
        return self.D if dim is not None else (self.D, self.D)

    def __len__(self):
        return self.L

    @property
    def order(self):
        return lib().edk_basis_order(self.h)

    def tables(self):
        """(I, R): 1-based representatives and norms, exactly as b.I, b.R."""
        D = self.D
        I = np.zeros(D, np.int64)
        R = np.zeros(D, np.float64)
        check(lib().edk_basis_download(self.h, I.ctypes.data, R.ctypes.data))
        return I, R

    @property
    def I(self):  # noqa: E743
        return self.tables()[0]

    @property
    def R(self):
        return self.tables()[1]

    def content(self, dgt):
        """index(dgt; base) - 1 (AbstractBasis.jl:285-298) of a digit row."""
        c = 0
        for d in dgt:
            c = c * self.base + int(d)
        return c

    def index(self, contents):
        """index(b, dgt) for 0-based contents: (coef, 1-based pos); absent -> (0, 1)."""
        s = np.ascontiguousarray(np.atleast_1d(np.asarray(contents, np.int64)))
        n = s.size
        pos = np.zeros(n, np.int64)
        cre = np.zeros(n)
        cim = np.zeros(n)
        check(lib().edk_basis_index(self.h, s.ctypes.data, n, pos.ctypes.data, cre.ctypes.data,
                                    cim.ctypes.data))
        return cre + 1j * cim, pos

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().edk_basis_destroy(self.h)
            except Exception:
                pass
            self.h = None


class TensorBasis(AbstractBasis):
    def __init__(self, L, base=2, ctx=None):
        super().__init__(_lib.EDK_TENSOR, L, base, ctx=ctx)


class ProjectedBasis(AbstractBasis):
    def __init__(self, L, N=None, f=None, base=2, ctx=None):
        if N is None:
            raise NotImplementedError("ProjectedBasis without N: pass N (f alone would scan base^L states)")
        super().__init__(_lib.EDK_PROJECTED, L, base, N=N, f=f, ctx=ctx)


class TranslationalBasis(AbstractBasis):
    def __init__(self, L, k=0, N=None, a=1, f=None, base=2, ctx=None):
        super().__init__(_lib.EDK_TRANSLATIONAL, L, base, a=a, N=N, k=k, f=f, ctx=ctx)
        ln = self.L // self.a
        kp = (-k) % ln
        self.complex_phases = not (kp == 0 or 2 * kp == ln)


class TranslationParityBasis(AbstractBasis):
    def __init__(self, L, k=0, p=1, N=None, a=1, f=None, base=2, ctx=None):
        super().__init__(_lib.EDK_TRANSLATION_PARITY, L, base, a=a, N=N, k=k, p=p, f=f, ctx=ctx)


class ParityBasis(AbstractBasis):
    """ParityBasis(; L, f, p, N, base): one representative per reflection orbit."""
    complex_phases = True  # eltype(::AbstractBasis) = ComplexF64 (AbstractBasis.jl:121)

    def __init__(self, L, p, N=None, f=None, base=2, ctx=None):
        super().__init__(_lib.EDK_PARITY, L, base, N=N, p=p, f=f, ctx=ctx)


class FlipBasis(AbstractBasis):
    """FlipBasis(; L, f, p, N, base): one representative per spin-flip orbit
    (digit d -> base - 1 - d); N must be L(base - 1)/2 when given."""
    complex_phases = True

    def __init__(self, L, p, N=None, f=None, base=2, ctx=None):
        super().__init__(_lib.EDK_FLIP, L, base, N=N, p=p, f=f, ctx=ctx)


class ParityFlipBasis(AbstractBasis):
    """ParityFlipBasis(; L, f, p, z, N, base): reflection and spin flip."""
    complex_phases = True

    def __init__(self, L, p, z, N=None, f=None, base=2, ctx=None):
        super().__init__(_lib.EDK_PARITY_FLIP, L, base, N=N, p=p, z=z, f=f, ctx=ctx)


class TranslationFlipBasis(AbstractBasis):
    """TranslationFlipBasis(; f, k, p, L, N, a, base): translation by the unit
    cell and spin flip with eigenvalue p."""

    def __init__(self, L, k=0, p=1, N=None, a=1, f=None, base=2, ctx=None):
        super().__init__(_lib.EDK_TRANSLATION_FLIP, L, base, a=a, N=N, k=k, p=p, f=f, ctx=ctx)
        ln = self.L // self.a
        kp = (-k) % ln
        # eltype follows phase_factor: real at k' in {0, len/2}
        self.complex_phases = not (kp == 0 or 2 * kp == ln)


class AbelianBasis(AbstractBasis):
    complex_phases = True  # eltype(::AbstractBasis) = ComplexF64 (AbstractBasis.jl:121)


def basis(L, N=None, k=None, a=1, p=None, z=None, symmetries=None, base=2, f=None, ctx=None):
    """basis(...) front door (AbelianBasis.jl:771-847)."""
    if symmetries is not None and (k is not None or p is not None or z is not None):
        raise ValueError("Cannot specify both `symmetries` and `k`/`p`/`z` keywords")
    if k is None and p is None and z is None and not symmetries:
        if f is None and N is None:
            return TensorBasis(L, base, ctx=ctx)
        return ProjectedBasis(L, N=N, f=f, base=base, ctx=ctx)
    # the Gosper path (base 2 with N) drops the predicate (AbelianBasis.jl:530-540)
    if base == 2 and N is not None:
        f = None
    b = AbelianBasis.__new__(AbelianBasis)
    AbstractBasis.__init__(b, _lib.EDK_ABELIAN, L, base, a=a, N=N, k=k or 0, p=p or 0, z=z or 0,
                           has_k=k is not None, has_p=p is not None, has_z=z is not None,
                           symmetries=symmetries, f=f, ctx=ctx)
    return b


def _host_tables(B):
    """b.I and b.R downloaded once per basis (the tables do not change)."""
    t = getattr(B, "_host_tab", None)
    if t is None:
        t = B.tables()
        B._host_tab = t
    return t


def _check_i(B, i):
    i = int(i)
    if not 1 <= i <= B.D:
        raise IndexError(f"basis index {i} outside 1..{B.D}")
    return i


def content(B, i):
    """content(b, i) (AbstractBasis.jl:95, :240): the stored label b.I[i] of the
    i-th (1-based) basis state, the representative's integer content plus one;
    i itself for TensorBasis."""
    i = _check_i(B, i)
    if B.kind == _lib.EDK_TENSOR:
        return i
    return int(_host_tables(B)[0][i - 1])


def norm(B, i):
    """norm(b, i) (AbstractBasis.jl:106-107, :241): b.R[i]; 1 for TensorBasis
    and ProjectedBasis."""
    i = _check_i(B, i)
    if B.kind in (_lib.EDK_TENSOR, _lib.EDK_PROJECTED):
        return 1.0
    return float(_host_tables(B)[1][i - 1])


def change_(B, i):
    """change!(b, i) (AbstractBasis.jl:160-163): the digits of the i-th basis
    state (site 1 first) and its norm.  The reference writes the digits into
    b.dgt and returns the norm; here both are returned."""
    return B.digits(content(B, i) - 1), norm(B, i)


def index(B, dgt):
    """index(b, dgt) (ProjectedBasis.jl:52-57, TranslationalBasis.jl:310-341,
    TranslationalParityBasis.jl:192-234, AbelianBasis.jl:669-693, ...): the
    coefficient and 1-based position of the digit row dgt (site 1 first) in B;
    (0, 1) when its representative is not in the sector."""
    dgt = np.asarray(dgt)
    if dgt.shape != (B.L,) or np.any(dgt < 0) or np.any(dgt >= B.base):
        raise ValueError(f"index: expected {B.L} digits in 0..{B.base - 1}")
    coef, pos = B.index([B.content(dgt)])
    c = complex(coef[0])
    return (c if c.imag != 0 else c.real), int(pos[0])


def productstate(v, B):
    """productstate (ToolKit.jl:103-109): real unit vector at index(B, v)."""
    coef, pos = B.index([B.content(v)])
    s = np.zeros(B.D)
    s[pos[0] - 1] = 1.0
    return s
