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


def test_candidates_any_length_with_rotation_counts():
    """The same argument on L-bit words, 16 <= L <= 32, with the rotation
    counts that fix the phase: the first (smallest) right-rotation count
    reaching the forward minimum, and the first count k with
    reverse(rotr^k(s)) at the reflected minimum, as the reference's loops
    find them; a word with at most two longest runs has period L or L/2."""
    rnd = random.Random(11)
    for L in range(16, 33):
        mask = (1 << L) - 1

        def rl(x, m):
            m %= L
            return ((x << m) | (x >> (L - m))) & mask if m else x

        def rr(x, m):
            return rl(x, (L - m) % L)

        def rv(x):
            return int(f"{x:0{L}b}"[::-1], 2)

        words = [rnd.getrandbits(L) for _ in range(300)]
        words += [sum(1 << p for p in rnd.sample(range(L), L // 2)) for _ in range(300)]
        words += [sum((rnd.getrandbits(L // 2)) << (k * (L // 2)) for k in range(2))
                  for _ in range(50)] if L % 2 == 0 else []
        for s in words:
            z = {1: ~s & mask}
            for m in (2, 4, 8, 16):
                z[m] = z[m // 2] & rl(z[m // 2], m // 2)
            acc, ln = mask, 0
            for m in (16, 8, 4, 2, 1):
                c = acc & rl(z[m], ln % L)
                if c:
                    acc, ln = c, ln + m
            cands = [p for p in range(L) if acc >> p & 1]
            if len(cands) > 2:
                continue
            fa = [(rr(s, (p + 1) % L), (p + 1) % L) for p in cands]
            fb = [(rl(rv(s), (p + 1 - ln) % L), (p + 1 - ln) % L) for p in cands]
            minA = min(v for v, _ in fa)
            Ma = min(k for v, k in fa if v == minA)
            minB = min(v for v, _ in fb)
            Mb = min(k for v, k in fb if v == minB)
            imgs = [rr(s, k) for k in range(L)]
            refl = [rv(rr(s, k)) for k in range(L)]
            assert (minA, Ma) == (min(imgs), imgs.index(min(imgs))), (L, bin(s))
            assert (minB, Mb) == (min(refl), refl.index(min(refl))), (L, bin(s))
            per = next(p for p in range(1, L + 1) if L % p == 0 and rl(s, p) == s)
            assert per == (L // 2 if L % 2 == 0 and rl(s, L // 2) == s else L), (L, bin(s))


def test_window_form_any_length():
    """The window form used on the device for 16 <= L < 32: an L-bit ring word
    x kept top-aligned in 32 bits with its low c = 32 - L bits continuing the
    ring, so every rotation is one funnel shift of (W, W << c).  Same
    representatives, rotation counts, reflected minima and periods as brute
    force, for every L and filling."""
    M32 = 0xffffffff
    rnd = random.Random(5)

    def alignbit(a, b, sh):
        return (((a << 32) | b) >> (sh & 31)) & M32

    def canon_w(s, L):
        c = 32 - L
        W = ((s << c) | (s >> (L - c))) & M32
        Wc = (W << c) & M32

        def wl(X, XC, m):  # rotl by 1 <= m <= L
            return alignbit(X, XC, 32 - m)

        def wr(X, XC, m):  # rotr by 0 <= m < L
            return alignbit(X, XC, c + m)
        z, zc = {1: ~W & M32}, {}
        zc[1] = (z[1] << c) & M32
        for m in (2, 4, 8, 16):
            z[m] = z[m // 2] & wl(z[m // 2], zc[m // 2], m // 2)
            zc[m] = (z[m] << c) & M32
        acc, ln = (z[16], 16) if z[16] else (M32, 0)
        for m in (8, 4, 2, 1):
            sh = ln - L if ln >= L else ln
            cc = acc & wl(z[m], zc[m], sh or L)
            if cc:
                acc, ln = cc, ln + m
        top = acc & (M32 << c) & M32
        cands = [b - c for b in range(c, 32) if top >> b & 1]
        if len(cands) > 2:
            return None
        fa = [(wr(W, Wc, (p + 1) % L), (p + 1) % L) for p in cands]
        b = brev(s)
        Wb = b | (b >> L)
        Wbc = (Wb << c) & M32
        fb = [(wl(Wb, Wbc, (p + 1 - ln) % L or L), (p + 1 - ln) % L) for p in cands]
        minA = min(v for v, _ in fa)
        minB = min(v for v, _ in fb)
        per = L // 2 if L % 2 == 0 and wl(W, Wc, L // 2) == W else L
        return (minA >> c, min(k for v, k in fa if v == minA), minB >> c,
                min(k for v, k in fb if v == minB), per)

    for L in range(16, 32):
        mask = (1 << L) - 1

        def rl(x, m):
            m %= L
            return ((x << m) | (x >> (L - m))) & mask if m else x

        def rv(x):
            return int(f"{x:0{L}b}"[::-1], 2)
        words = [rnd.getrandbits(L) for _ in range(150)]
        words += [sum(1 << p for p in rnd.sample(range(L), n)) for n in range(1, L) for _ in range(6)]
        if L % 2 == 0:
            words += [sum(rnd.getrandbits(L // 2) << (k * (L // 2)) for k in range(2)) for _ in range(40)]
        for s in words:
            r = canon_w(s, L)
            if r is None:
                continue
            imgs = [rl(s, (L - k) % L) for k in range(L)]
            refl = [rv(v) for v in imgs]
            per = next(p for p in range(1, L + 1) if L % p == 0 and rl(s, p) == s)
            assert r[:4] == (min(imgs), imgs.index(min(imgs)), min(refl), refl.index(min(refl))), (L, bin(s))
            assert r[4] == (L // 2 if L % 2 == 0 and (L // 2) % per == 0 else L), (L, bin(s))


def test_window_full_search_any_length():
    """The deferred words' full search on periodic windows (16 <= L < 32): the
    first strict minimum over increasing rotation counts is the reference's
    phase index, the first count mapping the word onto itself its period;
    checked against brute force on random and periodic words."""
    M32 = 0xffffffff
    rnd = random.Random(9)

    def alignbit(a, b, sh):
        return (((a << 32) | b) >> (sh & 31)) & M32

    for L in range(16, 32):
        mask, c = (1 << L) - 1, 32 - L

        def rl(x, m):
            m %= L
            return ((x << m) | (x >> (L - m))) & mask if m else x

        def rv(x):
            return int(f"{x:0{L}b}"[::-1], 2)
        words = [rnd.getrandbits(L) for _ in range(100)]
        for p in range(1, L):
            if L % p == 0:
                for _ in range(10):
                    u = rnd.getrandbits(p)
                    words.append(sum(u << (k * p) for k in range(L // p)))
        for s in words:
            W = ((s << c) | (s >> (L - c))) & M32
            Wc = (W << c) & M32
            minA, Ma, per = W, 0, L
            for m in range(1, L):
                v = alignbit(W, Wc, c + m)
                Ma = m if v < minA else Ma
                minA = min(minA, v)
                per = m if (v == W and per == L) else per
            b = brev(s)
            Wb = b | (b >> L)
            Wbc = (Wb << c) & M32
            minB, Mb = Wb, 0
            for k in range(1, L):
                v = alignbit(Wb, Wbc, 32 - k)
                Mb = k if v < minB else Mb
                minB = min(minB, v)
            imgs = [rl(s, (L - k) % L) for k in range(L)]
            refl = [rv(x) for x in imgs]
            assert (minA >> c, Ma) == (min(imgs), imgs.index(min(imgs))), (L, bin(s))
            assert (minB >> c, Mb) == (min(refl), refl.index(min(refl))), (L, bin(s))
            assert per == next(p for p in range(1, L + 1) if L % p == 0 and rl(s, p) == s), (L, bin(s))
