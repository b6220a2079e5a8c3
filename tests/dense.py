"""Independent dense constructions used to pin the oracle (small L only).

Nothing here restates the reference's enumeration or index() code: sectors
are defined by group theory (orbits, characters, the unnormalised projector
sum_g chi(g) g), operators by explicit Kronecker embedding.
"""
import itertools
from math import comb, gcd

import numpy as np


def digits(c, L, base):
    d = []
    for _ in range(L):
        c, r = divmod(c, base)
        d.append(r)
    return d[::-1]


def content(d, base):
    c = 0
    for x in d:
        c = c * base + x
    return c


def kron_embed(mat, inds, L, base=2):
    """Full-space matrix of a local term (the construction of TestHelpers.jl:25-46)."""
    n = base ** L
    full = np.zeros((n, n), np.complex128)
    k = len(inds)
    for col in range(n):
        d2 = digits(col, L, base)
        j = content([d2[i - 1] for i in inds], base)
        for rl in range(base ** k):
            v = mat[rl, j]
            if v == 0:
                continue
            d1 = list(d2)
            for q, i in enumerate(inds):
                d1[i - 1] = digits(rl, k, base)[q]
            full[content(d1, base), col] += v
    return full


def full_operator(H, L, base=2):
    out = 0
    for m, ind in zip(H.M, H.I):
        out = out + kron_embed(m, ind, L, base)
    return out


def perm_action(d, perm):
    """Content of site i moves to site perm[i] (1-based perm)."""
    out = [0] * len(d)
    for i, p in enumerate(perm):
        out[p - 1] = d[i]
    return out


def group_elements(L, gens):
    """All elements of the abelian group generated by gens = [(perm, inv, order, phases)],
    in the odometer order (generator 1 fastest), as (site map function, phase)."""
    orders = [g[2] for g in gens]
    els = []
    for idx in itertools.product(*[range(o) for o in orders[::-1]]):
        idx = idx[::-1]
        ph = 1.0 + 0j
        for g, e in zip(gens, idx):
            ph *= g[3][e]
        els.append((idx, ph))
    return els


def apply_element(d, gens, idx, base):
    for (perm, inv, order, _), e in zip(gens, idx):
        for _ in range(e):
            d = perm_action(d, perm)
            if inv is not None:
                d = [base - 1 - x if f else x for x, f in zip(d, inv)]
    return d


def translation_gens(L, k, a=1):
    """Translation by a that moves site content i -> i - a (the basis(...) T),
    character exp(+2 pi i k m / (L/a))."""
    order = L // a
    perm = [((i - a - 1) % L) + 1 for i in range(1, L + 1)]
    ph = [np.exp(2j * np.pi * k * m / order) for m in range(order)]
    return (perm, None, order, ph)


def parity_gen(L, p):
    perm = [L - i + 1 for i in range(1, L + 1)]
    return (perm, None, 2, [1.0, float(p)])


def flip_gen(L, z):
    return (list(range(1, L + 1)), [1] * L, 2, [1.0, float(z)])


def sector_by_definition(L, base, gens, allowed):
    """Representatives (orbit minima, 0-based contents) whose projected vector
    sum_g chi(g) g|s> is nonzero, with R = ||sum_g chi(g) g|s>||."""
    els = group_elements(L, gens)
    seen = set()
    reps, norms = [], []
    for c in range(base ** L):
        d = digits(c, L, base)
        if not allowed(d):
            continue
        imgs = [content(apply_element(d, gens, idx, base), base) for idx, _ in els]
        if min(imgs) != c:
            continue
        vec = {}
        for (idx, ph), im in zip(els, imgs):
            vec[im] = vec.get(im, 0) + ph
        nrm = np.sqrt(sum(abs(v) ** 2 for v in vec.values()))
        if nrm > 1e-9:
            reps.append(c)
            norms.append(nrm)
    return np.array(reps, np.int64), np.array(norms)


def mobius(n):
    r, m, p = 1, n, 2
    while p * p <= m:
        if m % p == 0:
            m //= p
            if m % p == 0:
                return 0
            r = -r
        p += 1
    return -r if m > 1 else r


def divisors(n):
    return [d for d in range(1, n + 1) if n % d == 0]


def primitive_binary(n, j):
    """Primitive (aperiodic) binary necklaces of length n with j ones."""
    if n == 0:
        return 0
    return sum(mobius(d) * comb(n // d, j // d) for d in divisors(gcd(n, j))) // n


def translational_dim(L, N1, k):
    """Number of binary orbits with N1 ones whose period n has n*k' = 0 mod L
    (k' = -k mod L): the closed form of the momentum-k sector dimension."""
    kp = (-k) % L
    tot = 0
    for n in divisors(L):
        if (N1 * n) % L:
            continue
        if (n * kp) % L:
            continue
        tot += primitive_binary(n, N1 * n // L)
    return tot


def necklaces_fixed(L, N1):
    return sum(comb(L // d, N1 // d) * phi(d) for d in divisors(gcd(L, N1))) // L


def phi(n):
    return sum(1 for i in range(1, n + 1) if gcd(i, n) == 1)


def bracelets_fixed(L, N1):
    """Binary bracelets (dihedral orbits) with N1 ones: Burnside."""
    rot = sum(comb(L // d, N1 // d) * phi(d) for d in divisors(gcd(L, N1)))
    # reflections
    if L % 2 == 1:
        refl = L * comb((L - 1) // 2, N1 // 2)
    else:
        half = L // 2
        # axes through two vertices: L/2 of them; through edges: L/2
        v = sum(comb(half - 1, (N1 - e) // 2) * comb(2, e) for e in range(3)
                if (N1 - e) % 2 == 0 and (N1 - e) >= 0)
        ed = comb(half, N1 // 2) if N1 % 2 == 0 else 0
        refl = half * v + half * ed
    return (rot + refl) // (2 * L)
