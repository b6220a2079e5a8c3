// edk_fast.h -- shared device code of the specialised gather kernels
// (edk_fast.hip: one vector; edk_fast_blk.hip: blocks of vectors).
//
// edk_fast.hip -- specialised gather kernel for translation-invariant
// two-site operators on 32-bit packed words.
//
// Applies when (checked on the host, else the generic kernel runs):
//   - base 2, L <= 32, so a configuration is one 32-bit word;
//   - the basis is TranslationParityBasis, TranslationalBasis (any k, unit
//     cell a) or ProjectedBasis, built by the library (orbit norms known);
//   - the operator is a sum of "families": for a fixed 2x2-site matrix M and
//     site distance d, the terms M on [i, i+d mod L] for every i -- exactly
//     what trans_inv_operator(M, [1, 1+d], B) builds (Operator.jl:102-108);
//   - M is real.
//
// Row i with configuration s: for a family (M, d) let r = rotl(s, d); bit q of
// (s, r) is the pair (site i, site i+d) of the term whose first site sits at
// bit q.  The diagonal is sum_ab M[ab,ab] * popcount(sel_a(s) & sel_b(r));
// every off-diagonal entry M[loc, c] acts at the bits of
// sel(loc_a, s) & sel(loc_b, r), flipping (loc ^ c) at (q, q - d).  Targets
// are processed eight at a time: canonicalise all eight, then issue their
// directory reads, their in-bucket binary searches and their x gathers
// together, so each lane keeps eight independent memory requests in flight.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "edk_matvec.h"

#define FAST_MAXFAM 4
#define FAST_MAXENT 12
#define FAST_TPB 256
#define FAST_TILE 4096

struct FamEnt {
  int locs;    // bit loc set: the entry acts where (first, second) site digits == loc
  int flip;    // bit 1: flip first site, bit 0: flip second site
  double val;  // M[loc, c], equal for every loc in locs
};

// positions (bits of the first site) whose local configuration is in locs
__device__ __forceinline__ uint32_t ent_mask(int locs, uint32_t s, uint32_t ns, uint32_t r, uint32_t mask) {
  uint32_t nr = ~r & mask, m = 0;
  if (locs & 1) m |= ns & nr;
  if (locs & 2) m |= ns & r;
  if (locs & 4) m |= s & nr;
  if (locs & 8) m |= s & r;
  return m;
}

struct FastFam {
  int d;
  int nent;
  double diag[4];  // M[ab, ab]
  FamEnt ent[FAST_MAXENT];
};

struct FastTab {
  int nfam;
  FastFam fam[FAST_MAXFAM];
};

struct FastArgs {
  int kind, L, A, len, kp, P;
  uint32_t mask;
  int ndiv;
  int divs[16];  // divisors of len below len, ascending (orbit period candidates)
  int batch;     // targets per lane handled together (4, 8 or 16)
  int scan;      // bucket width searched with one wide scan (0: binary search only)
  const FastTab *tab;
  const double *Rtab;  // orbit norms (same table the enumerator used)
  const double2 *ph;   // TranslationalBasis phases C[m]
  const uint32_t *keys;  // D keys + padding of 0xffffffff
  const uint32_t *dir;
  int dshift;
  uint32_t ndir;
  uint32_t D;
  const double *Rrow;
  const uint2 *region;   // two-level directory: per region (key >> rshift) {base index, fine offset << 5 | shift}
  int rshift;            // region width in key bits
  const uint16_t *fine;  // per region: positions relative to base, one slot per 2^shift key values
  const uint32_t *chunk; // balanced schedule: first row of each chunk (nchunk + 1 entries), or null
  uint32_t nchunk;
  uint32_t xcdg;         // balanced schedule: consecutive chunks given to one XCD (0: plain order)
  // rank directory (null when it does not apply): every representative is a
  // word with the bits of fix1 set, those of fix0 clear and rn of the rbits
  // bits in dmask (from bit rlo up) set; its colex rank among such words
  // indexes granules of 64 ranks {representatives below the granule, bitmap}
  const uint4 *gran;
  uint32_t ngran;
  const uint32_t *rtab;   // 3-bit group table of the colex rank (RK_NT entries)
  uint32_t fix1, fix0, dmask;
  int rlo, rbits, rn, rng;  // domain: lowest bit, width, set bits; groups of 3 bits
};

__device__ __forceinline__ uint32_t rotr32(uint32_t s, int sh, int L, uint32_t mask) {
  if (L == 32) return __builtin_amdgcn_alignbit(s, s, sh);
  uint32_t hi = s >> (32 - L), lo = s | (s << L);
  return __builtin_amdgcn_alignbit(hi, lo, sh) & mask;
}
// the same for a length known to be below 32 (no branch on L)
__device__ __forceinline__ uint32_t rotr_short(uint32_t s, int sh, int L, uint32_t mask) {
  return __builtin_amdgcn_alignbit(s >> (32 - L), s | (s << L), sh) & mask;
}
// ring rotation for a kernel flavour: LT == 32 one v_alignbit, LT == 1
// (16 <= L < 32) a funnel without a branch, otherwise either
template <int LT>
__device__ __forceinline__ uint32_t rotr_lt(uint32_t s, int sh, int L, uint32_t mask) {
  if constexpr (LT == 32) return __builtin_amdgcn_alignbit(s, s, sh);
  else if constexpr (LT == 1) return rotr_short(s, sh, L, mask);
  else return rotr32(s, sh, L, mask);
}

__device__ __forceinline__ uint32_t rev32(uint32_t s, int L) { return __brev(s) >> (32 - L); }

// Canonicalise s'.  Returns false when the representative cannot be in the
// table; otherwise rep and the real (cr) or complex (cr, ci) factor
// conj(phase) * R_rep.  All loops have uniform trip counts (no per-lane
// divergence): the minimum over a whole cycle of len rotations equals the
// minimum over one period, and the first index attaining it lies in the
// first period, as in the reference's loops that stop at the period.
template <int KIND, int PH>
__device__ __forceinline__ bool fast_canon(const FastArgs &A, uint32_t s, uint32_t &rep, double &cr,
                                           double &ci) {
  if (KIND == EDK_PROJECTED) {
    rep = s;
    cr = 1.0;
    ci = 0.0;
    return true;
  }
  const int sh = A.A;
  int period = A.len;
  // constant indices into the argument struct: a run-time index would make
  // the compiler copy the whole struct to per-lane scratch at kernel entry
#pragma unroll
  for (int q = 0; q < 16; q++) {
    if (q < A.ndiv) {
      const int p = A.divs[q];
      const bool m = rotr32(s, p * sh, A.L, A.mask) == s;
      period = (m && period == A.len) ? p : period;
    }
  }
  if (KIND == EDK_TRANSLATIONAL) {
    uint32_t minA = s, cur = s;
    for (int i = 1; i < A.len; i++) {
      cur = rotr32(cur, sh, A.L, A.mask);
      minA = min(minA, cur);
    }
    rep = minA;
    if (((int64_t)period * A.kp) % A.len) return false;  // check_momentum
    double R = A.Rtab[period - 1];
    if (PH != 0) {
      int M = 0;
      cur = s;
      for (int i = 1; i < A.len; i++) {
        cur = rotr32(cur, sh, A.L, A.mask);
        M = (cur == minA && M == 0 && s != minA) ? i : M;
      }
      double2 c = A.ph[M];
      cr = c.x * R;
      ci = -c.y * R;
    } else {
      cr = R;
      ci = 0.0;
    }
    return true;
  }
  // TranslationParityBasis
  uint32_t rb = rev32(s, A.L);
  uint32_t minA = s, minB = rb, cur = s, curB = rb;
  for (int i = 1; i < A.len; i++) {
    cur = rotr32(cur, sh, A.L, A.mask);
    curB = rotr32(curB, sh, A.L, A.mask);
    minA = min(minA, cur);
    minB = min(minB, curB);
  }
  bool qp = (minA == minB);
  double R = A.Rtab[(qp ? A.len : 0) + period - 1];
  double sgn = 1.0;
  if (PH != 0) {
    // first index M with rot^M(s) (or reverse(rot^M(s))) at the minimum
    // (TranslationalParityBasis.jl:207-231)
    int Ma = -1, Mb = -1;
    cur = s;
    for (int i = 0; i < A.len; i++) {
      Ma = (cur == minA && Ma < 0) ? i : Ma;
      Mb = (rev32(cur, A.L) == minB && Mb < 0) ? i : Mb;
      cur = rotr32(cur, sh, A.L, A.mask);
    }
    if (minB < minA) sgn = (double)A.P * ((A.kp == 1 || (Mb & 1) == 0) ? 1.0 : -1.0);
    else sgn = (A.kp == 1 || (Ma & 1) == 0) ? 1.0 : -1.0;
  }
  rep = minA < minB ? minA : minB;
  cr = sgn * R;
  ci = 0.0;
  return true;
}

// The same canonicaliser for L = 32 and unit cell 1, fully unrolled: each of
// the 31 rotations is one v_alignbit, the minima fold into v_min3, the
// orbit period (a power of two dividing 32) is read off the same rotations.
template <int KIND, int PH>
__device__ __forceinline__ bool fast_canon32(const FastArgs &A, uint32_t s, uint32_t &rep, double &cr,
                                             double &ci, int *ri = nullptr) {
  if (KIND == EDK_PROJECTED) {
    rep = s;
    cr = 1.0;
    ci = 0.0;
    return true;
  }
  // rotations as a dependent chain (rotate by one each step): few live
  // registers, so the batch of targets can be interleaved without spills
  uint32_t minA = s, cur = s;
#pragma unroll
  for (int k = 0; k < 15; k++) {  // rotations 1..30 in pairs (v_min3)
    uint32_t c1 = __builtin_amdgcn_alignbit(cur, cur, 1), c2 = __builtin_amdgcn_alignbit(cur, cur, 2);
    minA = min(minA, min(c1, c2));
    cur = c2;
  }
  minA = min(minA, __builtin_amdgcn_alignbit(cur, cur, 1));
  int period = 32;
#pragma unroll
  for (int k = 16; k >= 1; k >>= 1) period = (__builtin_amdgcn_alignbit(s, s, k) == s) ? k : period;
  if (KIND == EDK_TRANSLATIONAL) {
    rep = minA;
    if (((int64_t)period * A.kp) % 32) return false;  // check_momentum
    double R = A.Rtab[period - 1];
    if (ri) *ri = period - 1;
    if (PH != 0) {
      int M = 0;
      bool found = (s == minA);
      cur = s;
#pragma unroll
      for (int k = 1; k < 32; k++) {
        cur = __builtin_amdgcn_alignbit(cur, cur, 1);
        M = (!found && cur == minA) ? k : M;
        found = found || cur == minA;
      }
      double2 c = A.ph[M];
      cr = c.x * R;
      ci = -c.y * R;
    } else {
      cr = R;
      ci = 0.0;
    }
    return true;
  }
  uint32_t rb = __brev(s);
  uint32_t minB = rb;
  cur = rb;
#pragma unroll
  for (int k = 0; k < 15; k++) {
    uint32_t c1 = __builtin_amdgcn_alignbit(cur, cur, 1), c2 = __builtin_amdgcn_alignbit(cur, cur, 2);
    minB = min(minB, min(c1, c2));
    cur = c2;
  }
  minB = min(minB, __builtin_amdgcn_alignbit(cur, cur, 1));
  bool qp = (minA == minB);
  double R = A.Rtab[(qp ? 32 : 0) + period - 1];
  if (ri) *ri = (qp ? 32 : 0) + period - 1;
  double sgn = 1.0;
  if (PH != 0) {
    // first k with rot^k(s) == minA, resp. reverse(rot^k(s)) == minB
    // (TranslationalParityBasis.jl:207-231)
    int Ma = 0, Mb = 0;
    bool fa = (s == minA), fb = (rb == minB);
    cur = s;
#pragma unroll
    for (int k = 1; k < 32; k++) {
      cur = __builtin_amdgcn_alignbit(cur, cur, 1);
      uint32_t cb = __brev(cur);
      Ma = (!fa && cur == minA) ? k : Ma;
      fa = fa || cur == minA;
      Mb = (!fb && cb == minB) ? k : Mb;
      fb = fb || cb == minB;
    }
    if (minB < minA) sgn = (double)A.P * ((A.kp == 1 || (Mb & 1) == 0) ? 1.0 : -1.0);
    else sgn = (A.kp == 1 || (Ma & 1) == 0) ? 1.0 : -1.0;
  }
  rep = minA < minB ? minA : minB;
  cr = sgn * R;
  ci = 0.0;
  return true;
}

// Run-length candidates for L = 32 (unit cell 1), phase mode 0.  With site 1
// in the top bit, the smallest rotation of a word starts at the beginning of
// one of its longest zero runs, so only those starts need comparing.  Binary
// lifting of the zero masks (z_2m = z_m & rotl(z_m, m)) gives the longest run
// length and the mask of its start positions; the reflected word has the same
// runs read backwards.  Takes up to two candidates (about 92 % of words at half
// filling); returns false for a word with more, which the caller hands to
// canon32_brute.  A word with at most two longest runs has period 32 or 16.
__device__ __forceinline__ bool canon32_runs(uint32_t s, uint32_t &minA, uint32_t &minB, int &period) {
  const uint32_t z1 = ~s;
  const uint32_t z2 = z1 & __builtin_amdgcn_alignbit(z1, z1, 31);
  const uint32_t z4 = z2 & __builtin_amdgcn_alignbit(z2, z2, 30);
  const uint32_t z8 = z4 & __builtin_amdgcn_alignbit(z4, z4, 28);
  const uint32_t z16 = z8 & __builtin_amdgcn_alignbit(z8, z8, 24);
  uint32_t acc = z16 ? z16 : 0xffffffffu;
  uint32_t len = z16 ? 16u : 0u;
#define EDK_RUNSTEP(ZM, M)                                                   \
  {                                                                          \
    const uint32_t c_ = acc & __builtin_amdgcn_alignbit(ZM, ZM, (32u - len) & 31u); \
    const bool t_ = c_ != 0;                                                 \
    acc = t_ ? c_ : acc;                                                     \
    len = t_ ? len + (M) : len;                                              \
  }
  EDK_RUNSTEP(z8, 8u)
  EDK_RUNSTEP(z4, 4u)
  EDK_RUNSTEP(z2, 2u)
  EDK_RUNSTEP(z1, 1u)
#undef EDK_RUNSTEP
  if (__popc(acc) > 2) return false;
  const uint32_t rb = __brev(s);
  const uint32_t p1 = __builtin_ctz(acc);  // acc != 0: it only takes non-zero values
  const uint32_t m2 = acc & (acc - 1);
  const uint32_t p2 = m2 ? __builtin_ctz(m2) : p1;
  const uint32_t va1 = __builtin_amdgcn_alignbit(s, s, (p1 + 1) & 31u);
  const uint32_t va2 = __builtin_amdgcn_alignbit(s, s, (p2 + 1) & 31u);
  const uint32_t vb1 = __builtin_amdgcn_alignbit(rb, rb, (31u - p1 + len) & 31u);
  const uint32_t vb2 = __builtin_amdgcn_alignbit(rb, rb, (31u - p2 + len) & 31u);
  minA = min(va1, va2);
  minB = min(vb1, vb2);
  period = (__builtin_amdgcn_alignbit(s, s, 16) == s) ? 16 : 32;
  return true;
}

// Run-length candidates for any L with 16 <= L <= 32 and unit cell 1, every
// phase mode: the same longest-zero-run argument on L-bit words (LT == 32:
// one v_alignbit per rotation; otherwise a funnel rotation in L bits).  The
// rotation count of the chosen candidate gives the phase (first minimum, as
// the reference's loops: ties between the two candidates of a period-L/2
// word take the smaller count).  Returns 1 with rep and the factor
// conj(phase) * R (cr, ci) exactly as fast_canon / fast_canon32 compute it,
// 0 when the representative cannot be in the sector, -1 when the word has
// more than two longest runs (the caller runs the full search).  Given ri,
// these canonicalisers (and fast_canon32, fast_canon_w) also store the index
// of R in the norm table there.
template <int LT>
__device__ __forceinline__ uint32_t rotr_w(uint32_t x, uint32_t m, int L, uint32_t mask) {
  if (LT == 32) return __builtin_amdgcn_alignbit(x, x, m);
  return rotr32(x, (int)m, L, mask);  // 0 <= m < L
}
template <int LT>
__device__ __forceinline__ uint32_t rotl_w(uint32_t x, uint32_t m, int L, uint32_t mask) {
  if (LT == 32) return __builtin_amdgcn_alignbit(x, x, (32u - m) & 31u);
  return rotr32(x, m == 0 ? 0 : L - (int)m, L, mask);  // 0 <= m < L
}
// The run-length canonicaliser for 16 <= L < 32 on periodic windows: an
// L-bit ring word x is held top-aligned in 32 bits with its low c = 32 - L
// bits continuing the ring (W = x << c | x >> (L - c)), so the window is the
// first 32 bits of the infinite string x x x ...  Then a rotation is one
// v_alignbit of (W, W << c):
//   rotl_L(x, m) -> alignbit(W, W << c, 32 - m)   for 1 <= m <= L,
//   rotr_L(x, m) -> alignbit(W, W << c, c + m)    for 0 <= m <  L,
// windows order as their words do, and bitwise operations keep windows
// periodic.  Same candidates, rotation counts and results as canon_runs.
template <int KIND, int PH>
__device__ __forceinline__ int canon_runs_w(const FastArgs &B, uint32_t s, uint32_t &rep, double &cr, double &ci,
                                            int *ri = nullptr) {
  const uint32_t L = (uint32_t)B.L, c = 32u - L;
  const uint32_t W = (s << c) | (s >> (L - c)), Wc = W << c;
#define EDK_WROTL(X, XC, M) __builtin_amdgcn_alignbit((X), (XC), 32u - (M))
#define EDK_WROTR(X, XC, M) __builtin_amdgcn_alignbit((X), (XC), c + (M))
  // longest zero runs: binary lifting z_2m = z_m & rotl(z_m, m)
  const uint32_t z1 = ~W, z1c = z1 << c;
  const uint32_t z2 = z1 & EDK_WROTL(z1, z1c, 1u), z2c = z2 << c;
  const uint32_t z4 = z2 & EDK_WROTL(z2, z2c, 2u), z4c = z4 << c;
  const uint32_t z8 = z4 & EDK_WROTL(z4, z4c, 4u), z8c = z8 << c;
  const uint32_t z16 = z8 & EDK_WROTL(z8, z8c, 8u);
  uint32_t acc = z16 ? z16 : 0xffffffffu;
  uint32_t len = z16 ? 16u : 0u;
#define EDK_WSTEP(ZM, ZMC, M)                                    \
  {                                                              \
    uint32_t sh_ = len >= L ? len - L : len;                     \
    sh_ = sh_ ? sh_ : L;                                         \
    const uint32_t c_ = acc & EDK_WROTL(ZM, ZMC, sh_);           \
    const bool t_ = c_ != 0;                                     \
    acc = t_ ? c_ : acc;                                         \
    len = t_ ? len + (M) : len;                                  \
  }
  EDK_WSTEP(z8, z8c, 8u)
  EDK_WSTEP(z4, z4c, 4u)
  EDK_WSTEP(z2, z2c, 2u)
  EDK_WSTEP(z1, z1c, 1u)
#undef EDK_WSTEP
  const uint32_t top = acc & (0xffffffffu << c);  // one bit per ring position
  if (__popc(top) > 2) return -1;
  const uint32_t p1 = __builtin_ctz(top) - c;     // top != 0: acc is never 0
  const uint32_t m2 = top & (top - 1);
  const uint32_t p2 = m2 ? __builtin_ctz(m2) - c : p1;
  // forward images: rotate right by ia = (p + 1) mod L
  const uint32_t ia1 = (p1 + 1 == L) ? 0u : p1 + 1, ia2 = (p2 + 1 == L) ? 0u : p2 + 1;
  const uint32_t va1 = EDK_WROTR(W, Wc, ia1), va2 = EDK_WROTR(W, Wc, ia2);
  const uint32_t minA = min(va1, va2);
  uint32_t period = L;
  if ((L & 1) == 0 && EDK_WROTL(W, Wc, L >> 1) == W) period = L >> 1;
  if (KIND == EDK_TRANSLATIONAL) {
    rep = minA >> c;
    // check_momentum: period * k' divisible by L, period in {L, L/2}
    if (period != L && (B.kp & 1)) return 0;
    const double R = B.Rtab[period - 1];
    if (ri) *ri = period - 1;
    if (PH != 0) {
      const uint32_t Ma = (va1 == minA && (va2 != minA || ia1 < ia2)) ? ia1 : ia2;
      const double2 cc = B.ph[Ma];
      cr = cc.x * R;
      ci = -cc.y * R;
    } else {
      cr = R;
      ci = 0.0;
    }
    return 1;
  }
  // TranslationParityBasis: reflected images rev(rot_right^k(s)) =
  // rotl^k(rev(s)), k = (p + 1 - len) mod L; rev(s) is brev32(s) in the top
  // L bits, its window adds the continuation below
  const uint32_t b = __brev(s);
  const uint32_t Wb = b | (b >> L), Wbc = Wb << c;
  int ib1 = (int)p1 + 1 - (int)len, ib2 = (int)p2 + 1 - (int)len;
  ib1 += ib1 < 0 ? (int)L : 0;
  ib1 += ib1 < 0 ? (int)L : 0;
  ib2 += ib2 < 0 ? (int)L : 0;
  ib2 += ib2 < 0 ? (int)L : 0;
  const uint32_t vb1 = EDK_WROTL(Wb, Wbc, ib1 ? (uint32_t)ib1 : L);
  const uint32_t vb2 = EDK_WROTL(Wb, Wbc, ib2 ? (uint32_t)ib2 : L);
  const uint32_t minB = min(vb1, vb2);
  const bool qp = (minA == minB);
  const double R = B.Rtab[(qp ? L : 0) + period - 1];
  if (ri) *ri = (qp ? L : 0) + period - 1;
  double sgn = 1.0;
  if (PH != 0) {
    const uint32_t Ma = (va1 == minA && (va2 != minA || ia1 < ia2)) ? ia1 : ia2;
    const int Mb = (vb1 == minB && (vb2 != minB || ib1 < ib2)) ? ib1 : ib2;
    if (minB < minA) sgn = (double)B.P * ((B.kp == 1 || (Mb & 1) == 0) ? 1.0 : -1.0);
    else sgn = (B.kp == 1 || (Ma & 1) == 0) ? 1.0 : -1.0;
  }
  rep = (minA < minB ? minA : minB) >> c;
  cr = sgn * R;
  ci = 0.0;
  return 1;
#undef EDK_WROTL
#undef EDK_WROTR
}

// Full search on periodic windows (16 <= L < 32, unit cell 1) for the words
// canon_runs_w defers: every image is one v_alignbit; the first strict
// minimum over increasing rotation counts gives the reference's phase index,
// and the first count that maps the word onto itself is its period.
template <int KIND, int PH>
__device__ __forceinline__ bool fast_canon_w(const FastArgs &B, uint32_t s, uint32_t &rep, double &cr,
                                             double &ci, int *ri = nullptr) {
  const uint32_t L = (uint32_t)B.L, c = 32u - L;
  const uint32_t W = (s << c) | (s >> (L - c)), Wc = W << c;
  uint32_t minA = W, Ma = 0, period = L;
  for (uint32_t m = 1; m < L; m++) {
    const uint32_t v = __builtin_amdgcn_alignbit(W, Wc, c + m);  // rotr_L(s, m)
    Ma = v < minA ? m : Ma;
    minA = min(minA, v);
    period = (v == W && period == L) ? m : period;
  }
  if (KIND == EDK_TRANSLATIONAL) {
    rep = minA >> c;
    if (((int64_t)period * B.kp) % (int64_t)L) return false;  // check_momentum
    const double R = B.Rtab[period - 1];
    if (ri) *ri = period - 1;
    if (PH != 0) {
      const double2 cc = B.ph[Ma];
      cr = cc.x * R;
      ci = -cc.y * R;
    } else {
      cr = R;
      ci = 0.0;
    }
    return true;
  }
  // reflected images rev(rotr^k(s)) = rotl^k(rev(s)), k = 0 .. L-1
  const uint32_t b = __brev(s);
  const uint32_t Wb = b | (b >> L), Wbc = Wb << c;
  uint32_t minB = Wb, Mb = 0;
  for (uint32_t k = 1; k < L; k++) {
    const uint32_t v = __builtin_amdgcn_alignbit(Wb, Wbc, 32u - k);  // rotl_L(rev(s), k)
    Mb = v < minB ? k : Mb;
    minB = min(minB, v);
  }
  const bool qp = (minA == minB);
  const double R = B.Rtab[(qp ? L : 0) + period - 1];
  if (ri) *ri = (qp ? L : 0) + period - 1;
  double sgn = 1.0;
  if (PH != 0) {
    if (minB < minA) sgn = (double)B.P * ((B.kp == 1 || (Mb & 1) == 0) ? 1.0 : -1.0);
    else sgn = (B.kp == 1 || (Ma & 1) == 0) ? 1.0 : -1.0;
  }
  rep = (minA < minB ? minA : minB) >> c;
  cr = sgn * R;
  ci = 0.0;
  return true;
}

template <int KIND, int PH, int LT>
__device__ __forceinline__ int canon_runs(const FastArgs &B, uint32_t s, uint32_t &rep, double &cr, double &ci,
                                          int *ri = nullptr) {
  if constexpr (LT == 1) return canon_runs_w<KIND, PH>(B, s, rep, cr, ci, ri);
  const int L = (LT == 32) ? 32 : B.L;
  const uint32_t mask = (LT == 32) ? 0xffffffffu : B.mask;
  const uint32_t z1 = ~s & mask;
  const uint32_t z2 = z1 & rotl_w<LT>(z1, 1, L, mask);
  const uint32_t z4 = z2 & rotl_w<LT>(z2, 2, L, mask);
  const uint32_t z8 = z4 & rotl_w<LT>(z4, 4, L, mask);
  const uint32_t z16 = z8 & rotl_w<LT>(z8, 8, L, mask);
  uint32_t acc = z16 ? z16 : mask;
  uint32_t len = z16 ? 16u : 0u;
#define EDK_RUNSTEP(ZM, M)                                                             \
  {                                                                                    \
    const uint32_t sh_ = (LT == 32 || len < (uint32_t)L) ? len : len - (uint32_t)L;    \
    const uint32_t c_ = acc & rotl_w<LT>(ZM, sh_, L, mask);                            \
    const bool t_ = c_ != 0;                                                           \
    acc = t_ ? c_ : acc;                                                               \
    len = t_ ? len + (M) : len;                                                        \
  }
  EDK_RUNSTEP(z8, 8u)
  EDK_RUNSTEP(z4, 4u)
  EDK_RUNSTEP(z2, 2u)
  EDK_RUNSTEP(z1, 1u)
#undef EDK_RUNSTEP
  if (__popc(acc) > 2) return -1;
  const uint32_t p1 = __builtin_ctz(acc);  // acc != 0: it only takes non-zero values
  const uint32_t m2 = acc & (acc - 1);
  const uint32_t p2 = m2 ? __builtin_ctz(m2) : p1;
  // forward images: rotate right by ia = (p + 1) mod L
  const uint32_t ia1 = (p1 + 1 == (uint32_t)L) ? 0u : p1 + 1, ia2 = (p2 + 1 == (uint32_t)L) ? 0u : p2 + 1;
  const uint32_t va1 = rotr_w<LT>(s, ia1, L, mask), va2 = rotr_w<LT>(s, ia2, L, mask);
  const uint32_t minA = min(va1, va2);
  int period = L;
  if ((L & 1) == 0 && rotl_w<LT>(s, (uint32_t)(L >> 1), L, mask) == s) period = L >> 1;
  if (KIND == EDK_TRANSLATIONAL) {
    rep = minA;
    // check_momentum: period * k' divisible by L, period in {L, L/2}
    if (period != L && (B.kp & 1)) return 0;
    const double R = B.Rtab[period - 1];
    if (ri) *ri = period - 1;
    if (PH != 0) {
      const uint32_t Ma = (va1 == minA && (va2 != minA || ia1 < ia2)) ? ia1 : ia2;
      const double2 c = B.ph[Ma];
      cr = c.x * R;
      ci = -c.y * R;
    } else {
      cr = R;
      ci = 0.0;
    }
    return 1;
  }
  // TranslationParityBasis: reflected images rev(rot_right^k(s)) = rotl^k(rev(s)),
  // k = (p + 1 - len) mod L
  const uint32_t rb = (LT == 32) ? __brev(s) : rev32(s, L);
  int ib1 = (int)p1 + 1 - (int)len, ib2 = (int)p2 + 1 - (int)len;
  ib1 += ib1 < 0 ? L : 0;
  ib1 += ib1 < 0 ? L : 0;
  ib2 += ib2 < 0 ? L : 0;
  ib2 += ib2 < 0 ? L : 0;
  const uint32_t vb1 = rotl_w<LT>(rb, (uint32_t)ib1, L, mask), vb2 = rotl_w<LT>(rb, (uint32_t)ib2, L, mask);
  const uint32_t minB = min(vb1, vb2);
  const bool qp = (minA == minB);
  const double R = B.Rtab[(qp ? L : 0) + period - 1];
  if (ri) *ri = (qp ? L : 0) + period - 1;
  double sgn = 1.0;
  if (PH != 0) {
    const uint32_t Ma = (va1 == minA && (va2 != minA || ia1 < ia2)) ? ia1 : ia2;
    const int Mb = (vb1 == minB && (vb2 != minB || ib1 < ib2)) ? ib1 : ib2;
    if (minB < minA) sgn = (double)B.P * ((B.kp == 1 || (Mb & 1) == 0) ? 1.0 : -1.0);
    else sgn = (B.kp == 1 || (Ma & 1) == 0) ? 1.0 : -1.0;
  }
  rep = minA < minB ? minA : minB;
  cr = sgn * R;
  ci = 0.0;
  return 1;
}

// Full search for phase mode 0 at L = 32: representative and the index of its
// norm in the staged norm table (period - 1, plus 32 when the reflection
// maps the orbit onto itself for TranslationParityBasis).
template <int KIND>
__device__ __forceinline__ void canon32_brute(uint32_t s, uint32_t &rep, uint32_t &code) {
  uint32_t minA = s, cur = s;
#pragma unroll
  for (int k = 0; k < 15; k++) {
    uint32_t c1 = __builtin_amdgcn_alignbit(cur, cur, 1), c2 = __builtin_amdgcn_alignbit(cur, cur, 2);
    minA = min(minA, min(c1, c2));
    cur = c2;
  }
  minA = min(minA, __builtin_amdgcn_alignbit(cur, cur, 1));
  int period = 32;
#pragma unroll
  for (int k = 16; k >= 1; k >>= 1) period = (__builtin_amdgcn_alignbit(s, s, k) == s) ? k : period;
  if (KIND == EDK_TRANSLATIONAL) {
    rep = minA;
    code = period - 1;
    return;
  }
  uint32_t rb = __brev(s), minB = rb;
  cur = rb;
#pragma unroll
  for (int k = 0; k < 15; k++) {
    uint32_t c1 = __builtin_amdgcn_alignbit(cur, cur, 1), c2 = __builtin_amdgcn_alignbit(cur, cur, 2);
    minB = min(minB, min(c1, c2));
    cur = c2;
  }
  minB = min(minB, __builtin_amdgcn_alignbit(cur, cur, 1));
  rep = min(minA, minB);
  code = (minA == minB ? 32 : 0) + period - 1;
}

// marks a target whose representative is not in the table
#define HASH_EMPTY 0xffffffffu

// Position of key in the sorted table, or -1.  The directory bounds the
// search to [lo, hi); a bucket of at most A.scan keys is searched with one
// set of independent 16-byte loads (one memory round), a larger one by binary
// search.
__device__ __forceinline__ int64_t fast_find(const FastArgs &A, uint32_t key, uint32_t lo, uint32_t hi) {
  if (hi <= lo) return -1;
  if (hi - lo <= (uint32_t)A.scan) {
    uint32_t base = lo & ~3u;
    const uint4 *p = (const uint4 *)(A.keys + base);
    int below = 0, eq = 0;
    int nv = (int)((hi - base + 3) >> 2);
#pragma unroll 4
    for (int v = 0; v < nv; v++) {
      uint4 w = p[v];
      uint32_t idx = base + 4 * v;
      below += (idx + 0 >= lo && idx + 0 < hi && w.x < key);
      below += (idx + 1 >= lo && idx + 1 < hi && w.y < key);
      below += (idx + 2 >= lo && idx + 2 < hi && w.z < key);
      below += (idx + 3 >= lo && idx + 3 < hi && w.w < key);
      eq += (idx + 0 >= lo && idx + 0 < hi && w.x == key);
      eq += (idx + 1 >= lo && idx + 1 < hi && w.y == key);
      eq += (idx + 2 >= lo && idx + 2 < hi && w.z == key);
      eq += (idx + 3 >= lo && idx + 3 < hi && w.w == key);
    }
    return eq ? (int64_t)lo + below : -1;
  }
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (A.keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return (lo < A.D && A.keys[lo] == key) ? (int64_t)lo : -1;
}

// Two-level directory: region r = key >> rshift gives the region's first table
// index and its fine directory (sized to the region's population); the fine
// bucket bounds the key to [lo, hi), usually at most four keys, which are read
// with independent loads in one round.
#define FINE_SCAN 4
__device__ __forceinline__ void region_bounds(const FastArgs &A, uint32_t key, uint32_t &lo, uint32_t &hi) {
  uint2 rg = A.region[key >> A.rshift];
  uint32_t sh = rg.y & 31, foff = rg.y >> 5;
  uint32_t fb = (key & ((1u << A.rshift) - 1)) >> sh;
  lo = rg.x + A.fine[foff + fb];
  hi = rg.x + A.fine[foff + fb + 1];
}

typedef uint32_t edk_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint32_t region_find(const FastArgs &A, uint32_t key, uint32_t lo, uint32_t hi) {
  if (hi <= lo) return HASH_EMPTY;
  if (hi - lo <= FINE_SCAN) {
    // the bucket's keys in one 16-byte load (dword aligned; the table is
    // padded, so reading past the bucket or the end is harmless)
    static_assert(FINE_SCAN == 4, "one 16-byte load covers the scan");
    edk_u32x4 kv;
    __builtin_memcpy(&kv, A.keys + lo, sizeof(kv));
    const uint32_t n = hi - lo;
    const uint32_t below = (kv.x < key) + (n > 1 && kv.y < key) + (n > 2 && kv.z < key) + (n > 3 && kv.w < key);
    const bool eq = kv.x == key || (n > 1 && kv.y == key) || (n > 2 && kv.z == key) || (n > 3 && kv.w == key);
    return eq ? lo + below : HASH_EMPTY;
  }
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    uint32_t v = A.keys[mid];
    if (v == key) return mid;
    if (v < key) lo = mid + 1;
    else hi = mid;
  }
  return HASH_EMPTY;
}

// Branch-free lookups of BAL_K targets at once, so each step's loads for all
// targets issue back to back instead of one target's load waiting on the
// previous one's branch.  Invalid targets read slot 0 / key 0 (always in
// range; the key table is padded) and are masked afterwards.
#define EDK_NEED_BS 0xfffffffeu
template <int K>
__device__ __forceinline__ void region_lookup_many(const FastArgs &A, const uint32_t (&rep)[K], const bool (&ok)[K],
                                                   uint32_t (&jj)[K]) {
  uint2 rg[K];
#pragma unroll
  for (int q = 0; q < K; q++) rg[q] = A.region[(ok[q] ? rep[q] : 0u) >> A.rshift];
  // Only the bucket's first index is read.  The table is sorted, so every
  // key from the bucket's start up to the target is below it and every key
  // after it (later buckets, the padding) above it: among the four keys read,
  // the count below the target is its offset when present.  When all four are
  // below, the target may lie further on and the bucket's end is read.
  uint32_t lo[K], fslot[K];
#pragma unroll
  for (int q = 0; q < K; q++) {
    const uint32_t key = ok[q] ? rep[q] : 0u;
    const uint32_t sh = rg[q].y & 31, foff = rg[q].y >> 5;
    fslot[q] = foff + ((key & ((1u << A.rshift) - 1)) >> sh);
    lo[q] = rg[q].x + A.fine[fslot[q]];
  }
  edk_u32x4 kv[K];
#pragma unroll
  for (int q = 0; q < K; q++) __builtin_memcpy(&kv[q], A.keys + lo[q], sizeof(edk_u32x4));
  bool any_bs = false;
#pragma unroll
  for (int q = 0; q < K; q++) {
    const uint32_t key = rep[q];
    const uint32_t below = (kv[q].x < key) + (kv[q].y < key) + (kv[q].z < key) + (kv[q].w < key);
    const bool eq = kv[q].x == key || kv[q].y == key || kv[q].z == key || kv[q].w == key;
    jj[q] = !ok[q] ? HASH_EMPTY : (eq ? lo[q] + below : (below == 4 ? EDK_NEED_BS : HASH_EMPTY));
    any_bs = any_bs || jj[q] == EDK_NEED_BS;
  }
  if (any_bs) {
    // buckets of more than FINE_SCAN keys (rare): binary search over the rest
#pragma unroll
    for (int q = 0; q < K; q++) {
      if (jj[q] == EDK_NEED_BS) {
        uint32_t a = lo[q] + FINE_SCAN, b = rg[q].x + A.fine[fslot[q] + 1], r = HASH_EMPTY;
        while (a < b) {
          const uint32_t mid = (a + b) >> 1, v = A.keys[mid];
          if (v == rep[q]) {
            r = mid;
            break;
          }
          if (v < rep[q]) a = mid + 1;
          else b = mid;
        }
        jj[q] = r;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Rank directory.  Representatives with a fixed number of set bits are
// numbered by their colex rank, which for words of one popcount is their
// numeric order: with p_1 < p_2 < ... the positions of the set domain bits,
// rank = sum_i C(p_i, i).  A granule covers 64 consecutive ranks: the number
// of representatives of lower rank and a 64-bit membership bitmap in a
// 16-byte slot, so a lookup is one 16-byte load at rank / 64 (a shift: 96
// ranks a granule fill the slot but take quarter-rate multiplies to split a
// rank, and the rank kernel is bound by VALU issue).  The representatives cluster in
// rank (L = 32, N = 16: 9.4 M of them fill 13 % of the granules below the
// largest rank), so the directory's hot part is a few MB against the 56 MB of
// a key table with a bucket directory.
//
// Ranks are summed over 3-bit groups of the domain word from a table
// T[g][c][v] = sum over the set bits b of v, the j-th of them (j = 0, 1, ..),
// of C(3 g + b, c + j + 1), where c counts the set bits below the group.
// ---------------------------------------------------------------------------
#define RK_NCNT 17                   // c = 0 .. 16 (at most 16 domain bits set)
#define RK_NGRP 11                   // 3-bit groups of a 32-bit domain
#define RK_NT (RK_NGRP * RK_NCNT * 8)
#define RK_GRAN 64u                  // ranks per granule

// colex rank of the domain word w (ng groups)
__device__ __forceinline__ uint32_t rank_full(const uint32_t *T, uint32_t w, int ng) {
  uint32_t r = 0, c = 0;
  for (int g = 0; g < ng; g++) {
    const uint32_t v = (w >> (3 * g)) & 7u;
    r += T[(g * RK_NCNT + c) * 8 + v];
    c += __popc(v);
  }
  return r;
}

// C(n, k) for n < 32, k <= 17 from the group table (a single set bit); the
// index arithmetic stays in full-rate 24-bit multiplies (32-bit ones issue at
// a quarter of the rate, and the rank kernel is bound by VALU issue)
__device__ __forceinline__ uint32_t rank_binom(const uint32_t *T, int n, int k) {
  const uint32_t un = (uint32_t)n;
  const uint32_t g = __umul24(un, 11u) >> 5;  // n / 3 for n < 32
  const uint32_t b = un - __umul24(g, 3u);
  const uint32_t idx = __umul24(g, (uint32_t)(RK_NCNT * 8)) + ((uint32_t)(k - 1) << 3) + (1u << b);
  return k > 0 ? T[k > 0 ? idx : 0u] : 1u;
}

// colex terms of the set bits of a window v of at most 3 bits at domain
// position a, with c domain bits set below it
__device__ __forceinline__ uint32_t rank_window(const uint32_t *T, uint32_t v, int a, int c) {
  uint32_t r = 0;
#pragma unroll
  for (int b = 0; b < 3; b++) {
    const bool on = (v >> b) & 1u;
    c += on;
    r += on ? rank_binom(T, a + b, c) : 0u;
  }
  return r;
}

// a word is in the rank domain: fixed bits as required, rn domain bits set
__device__ __forceinline__ bool rank_domain(const FastArgs &A, uint32_t r) {
  return (r & A.fix1) == A.fix1 && (r & A.fix0) == 0 && __popc(r & A.dmask) == A.rn;
}

// positions in the table of K ranks (HASH_EMPTY: not a representative); the
// K granule loads issue together and every result is formed without a branch
template <int K>
__device__ __forceinline__ void rank_lookup_many(const FastArgs &A, const uint32_t (&rk)[K], const bool (&ok)[K],
                                                 uint32_t (&jj)[K]) {
  static_assert(RK_GRAN == 64u, "granules of two 32-bit bitmap words");
  uint4 g[K];
  bool in[K];
#pragma unroll
  for (int q = 0; q < K; q++) {
    const uint32_t gi = rk[q] / RK_GRAN;
    in[q] = ok[q] && gi < A.ngran;
    g[q] = A.gran[in[q] ? gi : 0u];
  }
#pragma unroll
  for (int q = 0; q < K; q++) {
    const uint32_t o = rk[q] % RK_GRAN, bit = o & 31u;
    const bool hi = o >= 32u;
    const uint32_t wsel = hi ? g[q].z : g[q].y;
    const uint32_t below = (hi ? __popc(g[q].y) : 0u) + __popc(wsel & ((1u << bit) - 1u));
    const bool hit = in[q] && ((wsel >> bit) & 1u);
    jj[q] = (g[q].x + below) | (hit ? 0u : HASH_EMPTY);
  }
}

// ---------------------------------------------------------------------------
// balanced kernel: one block per chunk of whole rows holding at most
// BAL_CAP targets.  Phase 1: the row threads write their targets (and entry
// numbers) into LDS in generator order.  Phase 2: every thread takes BAL_K
// targets (j = tid + BAL_TPB q), canonicalises them, looks them up and gathers
// x, all BAL_K independent so they overlap.  Phase 3: each row thread sums its
// contiguous slice of contributions in order and writes y.
// ---------------------------------------------------------------------------
// Chunk of workgroup b.  Workgroups are dispatched to the 8 XCDs round
// robin, so in plain order neighbouring chunks, whose rows share many
// targets, land on different L2 caches.  With G = xcdg, XCD x takes chunks in
// runs of G: its k-th workgroup gets chunk (k / G) 8G + x G + k % G.  All XCDs
// still advance through the rows together.  A bijection on the first
// multiple of 8G chunks; the tail keeps plain order.
__device__ __forceinline__ uint32_t bal_chunk(uint32_t b, uint32_t nchunk, uint32_t G) {
  if (G <= 1) return b;
  const uint32_t span = 8u * G, nfull = nchunk / span * span;
  if (b >= nfull) return b;
  const uint32_t x = b & 7u, k = b >> 3;
  return (k / G) * span + x * G + (k % G);
}

#define BAL_TPB 256
#define BAL_K 4
#define BAL_CAP (BAL_TPB * BAL_K)

__device__ __forceinline__ uint32_t block_scan_u32(uint32_t v, uint32_t *total, uint32_t *wsum) {
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t xx = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t yv = __shfl_up(xx, o, 64);
    if (lane >= o) xx += yv;
  }
  if (lane == 63) wsum[wid] = xx;
  __syncthreads();
  uint32_t base = 0, tot = 0;
  for (int q = 0; q < BAL_TPB / 64; q++) {
    if (q < wid) base += wsum[q];
    tot += wsum[q];
  }
  *total = tot;
  return base + xx - v;
}

// ---------------------------------------------------------------------------
// host: eligibility and setup (cached on the operator)
// ---------------------------------------------------------------------------
// Values of one operator for the multi-coefficient block kernel: the entry
// values and diagonals of its families (the structure -- distances, location
// masks, flips -- is the group's, identical for every column).
struct FastColTab {
  double val[FAST_MAXFAM * FAST_MAXENT];
  double diag[FAST_MAXFAM * 4];
};
#define MC_MAXV 8
// Per-column arguments of k_matvec_bal_mc: column c applies the values of
// ctab[c] to column c of the row-interleaved block and writes y[c] (a plain
// vector of D elements; null: a padding column, not written).
struct McArgs {
  const FastColTab *ctab[MC_MAXV];
  void *y[MC_MAXV];
};
struct ColPtrs {
  const void *p[MC_MAXV];
};

struct FastState {
  FastArgs args;
  FastTab *d_tab = nullptr;
  FastColTab *d_coltab = nullptr;  // this operator's values (multi-coefficient kernel)
  uint64_t signature = 0;          // hash of the family structure (distances, masks, flips)
  uint32_t *d_keys32 = nullptr;
  uint32_t *d_dir = nullptr;
  uint2 *d_region = nullptr;
  uint16_t *d_fine = nullptr;
  uint32_t *d_chunk = nullptr;
  uint4 *d_gran = nullptr;      // rank directory granules
  uint32_t *d_rtab = nullptr;   // its rank table
  bool cplx = false;
  int tuned[2][9] = {};  // schedule per (dtype, block width 1..8), EDK_VARIANT_*
};


// rank directory over sorted 32-bit representatives (edk_fast.hip); A.gran
// stays null when it does not apply; false only on a device error
bool edk_rank_build(edk_ctx *ctx, int L, const std::vector<uint32_t> &keys, const uint32_t *d_keys32, FastArgs &A,
                    uint4 **d_gran, uint32_t **d_rtab);

// block-of-vectors launch (edk_fast_blk.hip); nv in {2, 4, 8}
bool edk_fast_launch_blk(const FastArgs &A, int dtype, void *y, const void *x, int nv, int acc, hipStream_t st);
// Sectors of one fused block launch, passed by value in the kernel
// arguments (pointers read from there are known to be global memory; read
// from a device array they would take flat loads): sector k's arguments, its
// row-interleaved x and y and its first workgroup (bstart[n] = total
// workgroups).  12 sectors keep the struct under 3.5 KB.
#define MS_MAX 12
struct MultiSec {
  int n;
  uint32_t bstart[MS_MAX + 1];
  const void *x[MS_MAX];
  void *y[MS_MAX];
  FastArgs a[MS_MAX];
};
// phase mode of a sector's block launch (-1: none); sectors fuse when their
// kind, mode, length class and unit cell agree
int edk_fast_blk_mode(const FastArgs &A);
// one launch over the M.n sectors of M (the template follows M.a[0])
bool edk_fast_launch_blk_multi(const MultiSec &M, int dtype, int nv, int acc, hipStream_t st);
// one-vector launch with the rank directory (edk_fast_rk.hip); false when it
// does not apply (no directory, complex phases, not a run-length canonicaliser)
bool edk_fast_launch_rk(const FastArgs &A, int dtype, void *y, const void *x, int acc, hipStream_t st);
// multi-coefficient block launch (edk_fast_blk.hip): nv in {1, 2, 4, 8}
// columns of the row-interleaved block X (D x nv), column c with the values of
// M.ctab[c], written to M.y[c]; A is any member's arguments (shared structure)
bool edk_fast_launch_mc(const FastArgs &A, const McArgs &M, int dtype, const void *X, int nv, hipStream_t st);
// X[i * nv + c] = (G.p[c])[i] for nv input vectors of D elements
int edk_gather_cols(const ColPtrs &G, int nv, int64_t D, int dtype, void *X, hipStream_t st);
