"""Local spin matrices: spin(...) of EDKit (Operator.jl:536-718).

Lower-case x, y, z are spin operators S = sigma/2 (for D = 2), '+'/'-' the
ladder operators, '1'/'I' the identity; upper-case X, Y, Z (D = 2 only) are
the Pauli matrices.  Strings are Kronecker products, site order left to right.
"""
import numpy as np


def _spin_coeff(D):
    # spin_coeff (Operator.jl:536)
    return np.array([np.sqrt(i * (D - i)) for i in range(1, D)], dtype=np.float64)


def _dict(c, D):
    # spin_dict (Operator.jl:617-641)
    co = _spin_coeff(D)
    if c == "+":
        m = np.zeros((D, D))
        m[np.arange(D - 1), np.arange(1, D)] = co
        return m
    if c == "-":
        m = np.zeros((D, D))
        m[np.arange(1, D), np.arange(D - 1)] = co
        return m
    if c in "xy":
        sp = np.zeros((D, D))
        sp[np.arange(D - 1), np.arange(1, D)] = co / 2
        return sp + sp.T if c == "x" else sp - sp.T
    if c == "z":
        J = (D - 1) / 2
        return np.diag(np.arange(J, -J - 1, -1.0)[:D])
    if c in "1I":
        return np.eye(D)
    if D != 2:
        raise ValueError(f"Invalid spin symbol: {c}.")
    if c == "X":
        return np.array([[0.0, 1.0], [1.0, 0.0]])
    if c == "Y":
        return np.array([[0.0, 1.0], [-1.0, 0.0]])
    if c == "Z":
        return np.array([[1.0, 0.0], [0.0, -1.0]])
    raise ValueError(f"Invalid spin symbol: {c}.")


def _spin_string(s, D):
    # spin(s::String; D) (Operator.jl:691-699): raw Kronecker product times the
    # sign that turns the real i*S^y factors back into S^y
    if not s:
        raise ValueError("Empty spin string.")
    ny = sum(1 for c in s if c == "y" or (D == 2 and c == "Y"))
    raw = _dict(s[0], D)
    for c in s[1:]:
        raw = np.kron(raw, _dict(c, D))
    sign = (-1) ** (ny // 2) if ny % 2 == 0 else (-1j) ** ny
    return (sign * raw).astype(np.complex128)


def spin(*args, D=2):
    """spin("xx"), spin(c, "xx") or spin((c1, "xx"), (c2, "zz"), ...)."""
    if len(args) == 1 and isinstance(args[0], str):
        return _spin_string(args[0], D)
    if len(args) == 2 and isinstance(args[1], str) and not isinstance(args[0], tuple):
        return args[0] * _spin_string(args[1], D)
    terms = args[0] if len(args) == 1 and isinstance(args[0], list) else args
    out = None
    for c, s in terms:
        m = c * _spin_string(s, D)
        out = m if out is None else out + m
    return out
