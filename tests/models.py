"""Model builders shared by the tests and the bench (host-only term assembly)."""
import numpy as np

from edkit_hip.operator import operator, trans_inv_operator
from edkit_hip.spin import spin


class Shape:
    """Stand-in basis carrying only L and base (for host-side term assembly
    in CPU tests that never touch the device)."""

    def __init__(self, L, base=2):
        self.L, self.base = L, base
        self.complex_phases = True


def xxz(B, delta=1.0, jxy=1.0):
    """XXZ ring: Jxy (S^x S^x + S^y S^y) + delta S^z S^z on [i, i+1]."""
    return trans_inv_operator(spin((jxy, "xx"), (jxy, "yy"), (delta, "zz")), [1, 2], B)


def j1j2(B, j2=0.5, j1=1.0):
    """J1-J2 Heisenberg ring: J1 S_i.S_{i+1} + J2 S_i.S_{i+2}."""
    h = spin((1.0, "xx"), (1.0, "yy"), (1.0, "zz"))
    return trans_inv_operator(j1 * h, [1, 2], B) + trans_inv_operator(j2 * h, [1, 3], B)


def blbq(B, beta=0.2):
    """Spin-1 bilinear-biquadratic chain: S_i.S_{i+1} + beta (S_i.S_{i+1})^2."""
    h = spin((1.0, "xx"), (1.0, "yy"), (1.0, "zz"), D=3)
    return trans_inv_operator(h + beta * (h @ h), [1, 2], B)


def hopping_pm(B, delta=0.5):
    """The reference's sector Krylov test term: S+S- + S-S+ + delta S^z S^z."""
    return trans_inv_operator(spin((1.0, "+-"), (1.0, "-+"), (delta, "zz")), [1, 2], B)


def rand_state(D, seed, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(D) + 1j * rng.standard_normal(D)
    v /= np.linalg.norm(v)
    return v.astype(dtype)


def torus_sites(Lx, Ly):
    """Site (x, y) of an Lx x Ly torus is 1 + x + Lx y (x fastest)."""
    return [(x, y) for y in range(Ly) for x in range(Lx)]


def torus_translations(Lx, Ly):
    """T_x and T_y as site permutations (content of site i moves to site
    perm[i]), the form of the 2D example of basis(...) (AbelianBasis.jl:754-760)."""
    sites = torus_sites(Lx, Ly)
    tx = [(x + 1) % Lx + Lx * y + 1 for (x, y) in sites]
    ty = [x + Lx * ((y + 1) % Ly) + 1 for (x, y) in sites]
    return tx, ty


def heisenberg_2d(B, Lx, Ly, j=1.0, delta=1.0):
    """Nearest-neighbour XXZ model on an Lx x Ly torus (one bond per site and
    direction; Lx, Ly >= 3 so no bond repeats)."""
    h = j * spin((1.0, "xx"), (1.0, "yy"), (delta, "zz"))
    bonds = []
    for (x, y) in torus_sites(Lx, Ly):
        s = 1 + x + Lx * y
        bonds.append([s, 1 + (x + 1) % Lx + Lx * y])
        bonds.append([s, 1 + x + Lx * ((y + 1) % Ly)])
    return operator([h] * len(bonds), bonds, B)
