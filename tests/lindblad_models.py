"""Dense open-chain models for the Lindblad tests and bench (site 1 is the
most significant tensor factor, as in the bases)."""
import numpy as np

SX = np.array([[0, 1], [1, 0]], complex)
SY = np.array([[0, -1j], [1j, 0]], complex)
SZ = np.array([[1, 0], [0, -1]], complex)


def site_op(op, i, L):
    """op on site i (0-based) of L qubits, identity elsewhere."""
    return np.kron(np.kron(np.eye(2 ** i), op), np.eye(2 ** (L - i - 1)))


def dissipative_xxz(L, J=1.0, delta=0.5, gamma=0.2, hz=0.3):
    """Open XXZ chain H = J sum (X X + Y Y + delta Z Z)/4 + hz/2 sum Z_i and
    decay jumps sqrt(gamma) s-_i = sqrt(gamma) |0><1| on every site."""
    d = 2 ** L
    H = np.zeros((d, d), complex)
    for i in range(L - 1):
        for P, c in ((SX, 1.0), (SY, 1.0), (SZ, delta)):
            H += 0.25 * J * c * site_op(P, i, L) @ site_op(P, i + 1, L)
    for i in range(L):
        H += 0.5 * hz * site_op(SZ, i, L)
    sm = np.array([[0, 0], [1, 0]], complex)
    jumps = [np.sqrt(gamma) * site_op(sm, i, L) for i in range(L)]
    return H, jumps
