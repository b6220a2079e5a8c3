"""Operator assembly: operator(...) / trans_inv_operator(...) (Operator.jl:24-157).

Host-side term bookkeeping, as in the reference: a list of local matrices
and 1-based site lists (first site = most significant digit), identical site
lists merged by summing, zero matrices dropped.  The matvec itself runs in
libedkit_hip (see matvec.py).
"""
import numpy as np


def find_base(a, b):
    """find_base (Operator.jl:151-157): local dimension d with d^b == a."""
    if b == 1:
        return a
    for i in range(2, a + 1):
        if i ** b == a:
            return i
    raise ValueError(f"Incompatible dimension: ({a}, {b}).")


class Operator:
    """Operator{Tv,Tb}(M, I, B) (Operator.jl:24-28)."""

    def __init__(self, mats, inds, B):
        self.M = [np.asarray(m, dtype=np.complex128) for m in mats]
        self.I = [list(map(int, i)) for i in inds]
        self.B = B
        self._dev = {}

    # -- reference algebra (Operator.jl:117-147) --------------------------
    def __rmul__(self, c):
        if np.isscalar(c):
            return operator([c * m for m in self.M], self.I, self.B)
        return NotImplemented

    def __mul__(self, other):
        if np.isscalar(other):
            return other * self
        from .matvec import apply
        return apply(self, other)

    def __matmul__(self, v):
        from .matvec import apply
        return apply(self, v)

    def __truediv__(self, c):
        return operator([m / c for m in self.M], self.I, self.B)

    def __add__(self, other):
        if other is None:
            return self
        return operator(self.M + other.M, self.I + other.I, self.B)

    __radd__ = __add__

    def __neg__(self):
        return Operator([-m for m in self.M], self.I, self.B)

    def __sub__(self, other):
        return self + (-other)

    def adjoint(self):
        return operator([m.conj().T for m in self.M], self.I, self.B)

    def __len__(self):
        return len(self.M)

    @property
    def shape(self):
        if hasattr(self.B, "B1"):  # DoubleBasis: size(B1) x size(B2)
            return self.B.shape
        D = self.B.D
        return (D, D)

    def __repr__(self):
        return f"Operator of size {self.shape} with {len(self)} terms."

    def eltype(self):
        """promote_type(Tv, eltype(B)) (Operator.jl:30)."""
        cplx = any(np.any(m.imag != 0) for m in self.M) or getattr(self.B, "complex_phases", True)
        return np.complex128 if cplx else np.float64

    def flatten(self):
        """Flattened Vector{SparseMatrixCSC} + site lists for the C ABI:
        (nterms, site_off, sites, colptr_off, colptr, nz_off, rowval, nzval)."""
        site_off, sites, colptr_off, colptr, nz_off, rowval, nzval = [0], [], [], [], [0], [], []
        for m, ind in zip(self.M, self.I):
            n = m.shape[0]
            sites += ind
            site_off.append(len(sites))
            colptr_off.append(len(colptr))
            cp = [1]
            for c in range(n):
                rows = np.nonzero(m[:, c])[0]
                rowval += (rows + 1).tolist()
                nzval += m[rows, c].tolist()
                cp.append(cp[-1] + len(rows))
            colptr += cp
            nz_off.append(len(rowval))
        nz = np.zeros(2 * len(nzval), np.float64)
        if nzval:
            z = np.asarray(nzval, np.complex128)
            nz[0::2], nz[1::2] = z.real, z.imag
        as64 = lambda a: np.ascontiguousarray(np.asarray(a, np.int64))  # noqa: E731
        return (len(self.M), as64(site_off), as64(sites), as64(colptr_off), as64(colptr),
                as64(nz_off), as64(rowval), nz)


def operator(mats, inds, B):
    """operator(mats, inds, B) (Operator.jl:56-78), plus the single-matrix and
    spin-string forms (:85-88, :727-741)."""
    from .basis import TensorBasis
    from .spin import spin
    if isinstance(mats, str):
        base = B.base if hasattr(B, "base") else 2
        mats = [spin(mats, D=base)]
        inds = [inds]
    elif isinstance(mats, np.ndarray) and mats.ndim == 2:
        mats, inds = [mats], [inds]
    if inds and np.isscalar(inds[0]):
        inds = [[i] for i in inds]
    if len(mats) != len(inds):
        raise ValueError(f"Numbers mismatch: {len(mats)} matrices and {len(inds)} indices.")
    if not hasattr(B, "L"):
        base = find_base(np.asarray(mats[0]).shape[0], len(inds[0]))
        B = TensorBasis(L=int(B), base=base)
    M, I = [], []
    for m, ind in zip(mats, inds):
        m = np.asarray(m, np.complex128)
        if not np.any(m):
            continue
        ind = [int(i) for i in ind]
        if ind in I:
            M[I.index(ind)] = M[I.index(ind)] + m
        else:
            I.append(ind)
            M.append(m.copy())
    return Operator(M, I, B)


def trans_inv_operator(mat, ind, B):
    """trans_inv_operator(mat, ind, B) (Operator.jl:102-116, :749-774): the
    local term on sites mod.(ind .+ i, L) .+ 1 for i = -1:L-2."""
    from .basis import TensorBasis
    from .spin import spin
    if isinstance(mat, str):
        base = B.base if hasattr(B, "base") else 2
        if ind is None:
            ind = len(mat)
        mat = spin(mat, D=base)
    if np.isscalar(ind):
        ind = list(range(1, int(ind) + 1))
    if not hasattr(B, "L"):
        base = find_base(np.asarray(mat).shape[0], len(ind))
        B = TensorBasis(L=int(B), base=base)
    L = B.L
    inds = [[(i0 + i) % L + 1 for i0 in ind] for i in range(-1, L - 1)]
    return operator([mat] * L, inds, B)
