"""CPU check of the run-length candidate argument used by the L = 32
canonicaliser: with site 1 as the top bit, the smallest rotation of a 32-bit
word begins at the start of one of its longest zero runs (found by binary
lifting of the zero masks), and the reflected word's smallest rotation begins
at the end of one of the same runs.  Checked against the brute-force minimum
over all 64 dihedral images, on random words, half-filled words, periodic
words and the all-zero / all-one words."""
import random

M = 0xffffffff


def rotr(x, s):
    s &= 31
    return ((x >> s) | (x << (32 - s))) & M if s else x


def rotl(x, m):
    return rotr(x, 32 - m)


def brev(x):
    return int(f"{x:032b}"[::-1], 2)


def runs(s):
    """(candidate mask, longest zero-run length) by binary lifting."""
    z = {1: ~s & M}
    for m in (2, 4, 8, 16):
        z[m] = z[m // 2] & rotl(z[m // 2], m // 2)
    acc, ln = M, 0
    for m in (16, 8, 4, 2, 1):
        c = acc & rotl(z[m], ln)
        if c:
            acc, ln = c, ln + m
    return acc, ln


def test_candidates_give_the_dihedral_minimum():
    rnd = random.Random(7)
    words = [rnd.getrandbits(32) for _ in range(3000)]
    words += [sum(1 << p for p in rnd.sample(range(32), 16)) for _ in range(3000)]
    for p in (1, 2, 4, 8, 16):
        for _ in range(200):
            u = rnd.getrandbits(p)
            words.append(sum(u << (k * p) for k in range(32 // p)))
    words += [0, M, 1, M ^ 1, 0x0000FFFF, 0x55555555]
    for s in words:
        acc, ln = runs(s)
        cands = [p for p in range(32) if acc >> p & 1]
        assert cands
        minA = min(rotr(s, p + 1) for p in cands)
        minB = min(rotr(brev(s), 31 - p + ln) for p in cands)
        assert minA == min(rotl(s, r) for r in range(32)), hex(s)
        assert minB == min(rotl(brev(s), r) for r in range(32)), hex(s)
        # at most two longest runs: the period is 32 or 16
        if len(cands) <= 2:
            per = next(q for q in (1, 2, 4, 8, 16, 32) if rotl(s, q % 32) == s)
            assert per == (16 if rotr(s, 16) == s else 32), hex(s)
