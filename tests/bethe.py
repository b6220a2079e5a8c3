"""Bethe-ansatz ground energy of the spin-1/2 Heisenberg ring (test infrastructure).

H = J sum_i S_i . S_{i+1} on an even ring of L sites, S = sigma/2, with
M = L/2 flipped spins.  The ground state (the unique singlet) has rapidities
lambda_j solving the logarithmic Bethe equations

    L atan(2 lambda_j) = pi I_j + sum_l atan(lambda_j - lambda_l),

with I_j = -(M-1)/2, ..., (M-1)/2, and energy

    E0 = J L / 4 - (J / 2) sum_j 1 / (lambda_j^2 + 1/4)

(H. Bethe, Z. Phys. 71, 205 (1931); the form of M. Karbach and G. Mueller,
Computers in Physics 11, 36 (1997)).  Nothing here shares code with the
engine or the oracle: it is an independent closed form for the ground energy
of the L = 32 roofline sector, pinned against dense diagonalisation at small L
in tests/test_bethe.py.
"""
import numpy as np


def heisenberg_ring_e0(L, J=1.0, iters=2000, tol=1e-15):
    if L % 2:
        raise ValueError("even L only")
    M = L // 2
    I = np.arange(M) - (M - 1) / 2.0
    # start from the non-interacting rapidities, then damped fixed-point steps
    lam = 0.5 * np.tan(np.pi * I / L)
    for _ in range(iters):
        phi = (np.pi * I + np.arctan(lam[:, None] - lam[None, :]).sum(axis=1)) / L
        new = 0.5 * np.tan(phi)
        if np.max(np.abs(new - lam)) < 1e-13:
            lam = new
            break
        lam = 0.5 * (lam + new)
    # Newton polish on F_j = L atan(2 l_j) - pi I_j - sum_l atan(l_j - l_l)
    for _ in range(50):
        d = lam[:, None] - lam[None, :]
        F = L * np.arctan(2 * lam) - np.pi * I - np.arctan(d).sum(axis=1)
        g = 1.0 / (1.0 + d * d)
        Jm = g.copy()                      # dF_j/dl_l = + 1/(1 + (l_j - l_l)^2), l != j
        np.fill_diagonal(Jm, 0.0)
        diag = 2 * L / (1 + 4 * lam * lam) - (g.sum(axis=1) - 1.0)
        Jm[np.diag_indices(M)] = diag
        step = np.linalg.solve(Jm, F)
        lam = lam - step
        if np.max(np.abs(step)) < tol:
            break
    return J * L / 4.0 - 0.5 * J * np.sum(1.0 / (lam * lam + 0.25))
