"""CPU restatement of LinearMap.jl (basis_embedding, DoubleBasis(v),
symmetrizer) and of the column scatter of an operator on a DoubleBasis
(Operator.jl:418-459 with change!/index of DoubleBasis, LinearMap.jl:115-126).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, always as the checker, never
as the product path.  Bases are oracle.Basis objects; states are contents
(0-based integers in base `base`, site 1 the most significant digit).
"""
import numpy as np


def orbit_order(ob):
    """orbit_order (LinearMap.jl:25): 1 for on-site bases, order(B) otherwise."""
    return 1 if ob.kind in ("tensor", "projected") else ob.order


def _full_index(ob):
    return ob.index(np.arange(ob.base ** ob.L, dtype=np.uint64))


def basis_embedding(ob):
    """basis_embedding(B) (LinearMap.jl:57-70): dim(full) x size(B), entries coeff / ord."""
    pos, coef = _full_index(ob)
    E = np.zeros((ob.base ** ob.L, ob.D), np.complex128)
    nz = coef != 0
    E[np.nonzero(nz)[0], pos[nz] - 1] = coef[nz] / orbit_order(ob)
    return E


def double_map(ob1, ob2, v):
    """(B::DoubleBasis)(v) (LinearMap.jl:144-172): the loop over the full
    space in order, out[pos1] += conj(coeff1/ord1) (coeff2/ord2) v[pos2]."""
    v = np.asarray(v, np.complex128)
    if v.shape[0] != ob2.D:
        raise ValueError("Expected a vector of length %d; got %d." % (ob2.D, v.shape[0]))
    p2, c2 = _full_index(ob2)
    p1, c1 = _full_index(ob1)
    keep = (c2 != 0) & (c1 != 0)
    w = np.conj(c1[keep] / orbit_order(ob1)) * (c2[keep] / orbit_order(ob2))
    out = np.zeros((ob1.D,) + v.shape[1:], np.complex128)
    contrib = w.reshape((-1,) + (1,) * (v.ndim - 1)) * v[p2[keep] - 1]
    np.add.at(out, p1[keep] - 1, contrib)
    return out


def symmetrizer(ob1, ob2):
    """symmetrizer(DoubleBasis(B1, B2)) (LinearMap.jl:183-185)."""
    return basis_embedding(ob1).conj().T @ basis_embedding(ob2)


def _digits(c, L, base):
    return [(c // base ** (L - q)) % base for q in range(1, L + 1)]


def _content(dg, base):
    c = 0
    for d in dg:
        c = c * base + d
    return c


def double_operator(ob1, ob2, mats, inds):
    """The dense size(B1) x size(B2) matrix of operator(mats, inds,
    DoubleBasis(B1, B2)) by the reference's column loop: for the B2
    representative a_j (norm r = norm(B2, j)), every term and every nonzero
    M[row, loc] of column loc = local index of a_j, the image b' with the
    term's sites set to row adds M[row, loc] coeff1 / r at index(B1, b')."""
    L, base = ob1.L, ob1.base
    if ob2.kind == "tensor":
        I2, R2 = np.arange(1, base ** L + 1), np.ones(base ** L)
    else:
        I2, R2 = ob2.tables()
        if ob2.kind == "projected":
            R2 = np.ones_like(R2)
    H = np.zeros((ob1.D, ob2.D), np.complex128)
    for j in range(ob2.D):
        dg = _digits(int(I2[j]) - 1, L, base)
        r = R2[j]
        for M, I in zip(mats, inds):
            M = np.asarray(M, np.complex128)
            loc = _content([dg[i - 1] for i in I], base)
            for row in np.nonzero(M[:, loc])[0]:
                nd = list(dg)
                rd = _digits(int(row), len(I), base)
                for q, i in enumerate(I):
                    nd[i - 1] = rd[q]
                pos, coef = ob1.index(np.array([_content(nd, base)], np.uint64))
                if coef[0] != 0:
                    H[pos[0] - 1, j] += M[row, loc] * coef[0] / r
    return H
