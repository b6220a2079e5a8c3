// edk_perm.hip -- one-vector gather kernel for Abelian bases of spin-1/2
// sites with site-permutation generators (basis(...; symmetries): torus
// translations, custom flips and the like), over the rank directory.
//
// Applies when (checked on the host, else the generic kernel runs):
//   - the basis is Abelian, base 2, L <= 32, built by the library, and every
//     representative has one popcount (the rank directory, edk_fast.h);
//   - one to three generators, none a Benes network, each of order <= 32 and
//     the group of order <= 64; the walk is pruned when the basis has the
//     table of least top blocks with at most 128 entries (top blocks of at
//     most 7 bits) and generator 0 has order <= 16 with every power in at
//     most four rotation groups;
//   - the operator's terms are two-site terms with real matrices that group
//     into at most four families: terms of one matrix whose first sites are
//     distinct.  A family is a map p -> pi(p) of bit positions on its first
//     sites; the partner word r (bit p of r = bit pi(p) of s) takes one shift
//     per displacement p - pi(p), at most four (one for a ring, two for the
//     rows or the columns of a torus), so a family's entries act on the bits
//     of sel(s) & sel(r) as in the ring kernels.
//
// Per target the group walk (the Abelian walk of AbelianBasis.jl:428-445,
// pruned by generator 0's least top block when the basis has that table)
// gives the representative, the element reaching it and the stabiliser
// order; the coefficient is the entry value times conj(phase) times R of the
// target's orbit, and the row's sum is divided by its own R
// (Operator.jl:417-465, the generic kernel's formula).  Positions come from
// the rank directory: the colex rank of the representative indexes a 16-byte
// granule.
//
// Schedule: the balanced chunks of the ring kernels (whole rows, at most
// 1024 targets and 256 rows per workgroup, XCD-grouped order):
//   phase 1  row threads write their targets (word and entry) to LDS;
//   phase 2a every thread canonicalises and ranks four targets;
//   phase 2b granule loads, then x gathers, four in flight per thread; the
//            contributions then take the place of the target arrays;
//   phase 3  row threads sum their contributions in generation order
//            (family, entry, bit), add the diagonal and write y.
#include <atomic>
#include <map>
#include <mutex>

#include "edk_fast.h"

#define PM_MAXFAM 4
#define PM_MAXGRP 4
#define PM_MAXORD 32   // order of one generator
#define PM_MAXGRPORD 64  // order of the group
#define PM_MAXPT 128   // entries of the pruned walk's table held in LDS
#define PM_MAXPOW 16   // order of generator 0 in the pruned walk
#define PMM_ABSENT 0x4000u

struct PermFam {
  uint32_t fmask;           // bits of the family's first sites
  int ng;                   // displacement groups of the partner map
  uint32_t gm[PM_MAXGRP];   // first-site bits of group g
  uint32_t gr[PM_MAXGRP];   // r |= rotl(s, gr) & gm
  int nent;
  double diag[4];           // M[ab, ab]
  FamEnt ent[FAST_MAXENT];  // off-diagonal entries, as in the ring families
  uint8_t partner[32];      // bit of the second site, per first-site bit
};

// A generator on one-bit sites: an optional reversal of the L bits, then
// OR_q rotl(s & m[q], rl[q]) (a rotation of the word moves no masked bit
// across bit 31, so it is the shift by the group's displacement), then the
// spin inversion.  Rotations are two groups, the shift-and-mask forms their
// own, the identity one.
struct PermGen {
  int rev, n;
  uint32_t m[EDK_MAXSHIFTS];
  uint32_t rl[EDK_MAXSHIFTS];
  uint32_t inv;
  int order;
};

// A power of generator 0 in the same form, read by each lane from LDS with
// its own exponent: groups past the power's own are empty masks.
struct __align__(16) PowGen {
  uint32_t m[EDK_MAXSHIFTS];
  uint32_t rl[EDK_MAXSHIFTS];
  uint32_t inv, pad[3];
};

struct PermTab {
  int nfam;
  PermFam fam[PM_MAXFAM];
  int ngen;
  PermGen gen[3];
  int pbits;                          // pruned walk's table: 2^pbits entries (0: the full walk)
  uint32_t rcp0;                      // ceil(2^16 / order of generator 0): e / n0 = (e rcp0) >> 16 for e < 64
  uint32_t rcp1;                      // the same for generator 1 (an element's exponents without a division)
  PowGen pow0[PM_MAXPOW];             // generator 0's powers (pruned walk)
  double R[PM_MAXGRPORD];             // norm by stabiliser order
  double2 ph[3 * PM_MAXORD];          // phase of generator g's m-th power
};

struct PermState {
  FastArgs A;  // rows, chunks and rank directory (kind EDK_ABELIAN)
  const uint2 *ptab = nullptr;  // the basis' pruned-walk table, or null
  PermTab *d_tab = nullptr;
  uint32_t *d_keys32 = nullptr, *d_chunk = nullptr, *d_rtab = nullptr;
  uint4 *d_gran = nullptr;
  bool real = true;  // every phase real
  int npw = 1;       // the most rotation groups of any power of generator 0
};

__device__ __forceinline__ uint32_t rotl32(uint32_t v, uint32_t k) { return (v << k) | (v >> ((32u - k) & 31u)); }

__device__ __forceinline__ uint32_t perm_partner(const PermFam &F, uint32_t s) {
  uint32_t r = 0;
  for (int g = 0; g < F.ng; g++) r |= rotl32(s, F.gr[g]) & F.gm[g];
  return r;
}

// PermGen in registers (uniform: scalar registers)
struct Gen32 {
  int rev, n;
  uint32_t m[EDK_MAXSHIFTS], rl[EDK_MAXSHIFTS];
  uint32_t inv;
};

__device__ __forceinline__ Gen32 perm_gen(const PermTab &T, int g) {
  Gen32 r;
  r.rev = T.gen[g].rev;
  r.n = T.gen[g].n;
  r.inv = T.gen[g].inv;
#pragma unroll
  for (int q = 0; q < EDK_MAXSHIFTS; q++) {
    r.m[q] = T.gen[g].m[q];
    r.rl[q] = T.gen[g].rl[q];
  }
  return r;
}

__device__ __forceinline__ uint32_t perm_apply(int L, const Gen32 &r, uint32_t s) {
  if (r.rev) s = __brev(s) >> (32 - L);
  uint32_t o = rotl32(s & r.m[0], r.rl[0]);
#pragma unroll
  for (int q = 1; q < EDK_MAXSHIFTS; q++)
    if (q < r.n) o |= rotl32(s & r.m[q], r.rl[q]);
  return o ^ r.inv;
}

// NPW: at least the most groups any power has (a compile-time bound: a
// uniform branch per group costs more than the group)
template <int NPW>
__device__ __forceinline__ uint32_t perm_apply_pow(const PowGen &g, uint32_t s) {
  uint32_t o = rotl32(s & g.m[0], g.rl[0]);
#pragma unroll
  for (int q = 1; q < NPW; q++) o |= rotl32(s & g.m[q], g.rl[q]);
  return o ^ g.inv;
}

// The Abelian walk over 32-bit words with the generators in registers:
// the representative, the element number e = a0 + n0 (a1 + n1 a2) of the
// first strict minimum (odometer order, generator 0 fastest) and the
// stabiliser order.  PRUNE: only images whose top block is the least over
// generator 0's powers are compared (pt: the table in LDS, entries packed as
// least block | mask of the powers reaching it << 8; the pruning of
// edk_canon_abelian_pruned); each candidate is one power of generator 0 (pw)
// applied to its outer image, and the images equal to the minimum are a
// coset of the stabiliser, so their number is its order.
template <bool PRUNE, int NPW>
__device__ __forceinline__ void perm_canon(int L, const Gen32 &g0, const Gen32 &g1, const Gen32 &g2, int n0,
                                           int n1, int n2, const uint32_t *pt, int sh, const PowGen *pw,
                                           uint32_t rcp0, uint32_t s, uint32_t &rep, int &be, int &cnt) {
  uint32_t best = s;
  be = 0;
  cnt = 0;
  if constexpr (!PRUNE) {
    int e = 0;
    uint32_t c2 = s;
    for (int a2 = 0; a2 < n2; a2++) {
      uint32_t c1 = c2;
      for (int a1 = 0; a1 < n1; a1++) {
        uint32_t c0 = c1;
        for (int a0 = 0; a0 < n0; a0++, e++) {
          cnt += c0 == s;
          if (c0 < best) {
            best = c0;
            be = e;
          }
          if (a0 + 1 < n0) c0 = perm_apply(L, g0, c0);
        }
        if (a1 + 1 < n1) c1 = perm_apply(L, g1, c1);
      }
      if (a2 + 1 < n2) c2 = perm_apply(L, g2, c2);
    }
  } else if (n1 * n2 <= EDK_PRUNE_OUT) {
    uint32_t co[EDK_PRUNE_OUT];
    uint32_t c1 = s, c2 = s;
    int a1 = 0;
#pragma unroll
    for (int i = 0; i < EDK_PRUNE_OUT; i++) {
      co[i] = c1;
      if (i + 1 < n1 * n2) {
        if (++a1 < n1) {
          c1 = perm_apply(L, g1, c1);
        } else {
          a1 = 0;
          c2 = perm_apply(L, g2, c2);
          c1 = c2;
        }
      }
    }
    uint32_t te[EDK_PRUNE_OUT];
#pragma unroll
    for (int i = 0; i < EDK_PRUNE_OUT; i++) te[i] = pt[i < n1 * n2 ? (co[i] >> sh) : 0u];
    uint32_t vmin = 0xffffffffu;
#pragma unroll
    for (int i = 0; i < EDK_PRUNE_OUT; i++) vmin = i < n1 * n2 ? min(vmin, te[i] & 0xffu) : vmin;
    // the candidates as one mask over element numbers e = a0 + n0 i (the
    // group has at most 64 elements), taken in increasing e: the odometer
    // order of the walk; each costs one power of generator 0
    uint64_t cm = 0;
#pragma unroll
    for (int i = 0; i < EDK_PRUNE_OUT; i++)
      if (i < n1 * n2 && (te[i] & 0xffu) == vmin) cm |= (uint64_t)(te[i] >> 8) << (i * n0);
    while (cm) {
      const int e = __ffsll((unsigned long long)cm) - 1;
      cm &= cm - 1;
      const int i = (int)(__umul24((uint32_t)e, rcp0) >> 16);  // e / n0 (e < 64: a 24-bit product)
      uint32_t c = co[0];
#pragma unroll
      for (int k = 1; k < EDK_PRUNE_OUT; k++) c = i == k ? co[k] : c;
      c = perm_apply_pow<NPW>(pw[e - (int)__umul24((uint32_t)i, (uint32_t)n0)], c);
      if (cnt == 0 || c < best) {
        best = c;
        be = e;
        cnt = 1;
      } else if (c == best) {
        cnt++;
      }
    }
  } else {
    uint32_t vmin = 0xffffffffu;
    uint32_t c2 = s;
    for (int a2 = 0; a2 < n2; a2++) {
      uint32_t c1 = c2;
      for (int a1 = 0; a1 < n1; a1++) {
        vmin = min(vmin, pt[c1 >> sh] & 0xffu);
        if (a1 + 1 < n1) c1 = perm_apply(L, g1, c1);
      }
      if (a2 + 1 < n2) c2 = perm_apply(L, g2, c2);
    }
    c2 = s;
    for (int a2 = 0; a2 < n2; a2++) {
      uint32_t c1 = c2;
      for (int a1 = 0; a1 < n1; a1++) {
        const uint32_t t = pt[c1 >> sh];
        if ((t & 0xffu) == vmin) {
          uint32_t mask = t >> 8;
          uint32_t c0 = c1;
          int a0 = 0;
          while (mask) {
            const int nx = __ffs(mask) - 1;
            mask &= mask - 1;
            for (; a0 < nx; a0++) c0 = perm_apply(L, g0, c0);
            const int e = a0 + n0 * (a1 + n1 * a2);
            if (cnt == 0 || c0 < best) {
              best = c0;
              be = e;
              cnt = 1;
            } else if (c0 == best) {
              cnt++;
            }
          }
        }
        if (a1 + 1 < n1) c1 = perm_apply(L, g1, c1);
      }
      if (a2 + 1 < n2) c2 = perm_apply(L, g2, c2);
    }
  }
  rep = best;
}

// PH 0: every phase real (coefficients real); PH 2: complex coefficients.
// NPW: rotation groups of generator 0's powers (pruned walk; 2 or 4).
template <int PH, bool PRUNE, int NPW, typename VT, typename ST>
__global__ void __launch_bounds__(BAL_TPB, PH == 0 ? 6 : 4)
    k_matvec_perm(FastArgs A, const PermTab *__restrict__ tab, const uint2 *__restrict__ ptab,
                  const VT *__restrict__ x, VT *__restrict__ y, int accumulate) {
  // per target: word (then rank), meta word (entry 0-5, absent 14, row in
  // the chunk 16-23) and coefficient, as three arrays; the contributions
  // overlay them after the last reads of phase 2b
  typedef typename std::conditional<PH == 0, ST, VT>::type CT;
  typedef typename std::conditional<PH == 0, double, double2>::type PT;
  constexpr int UB = BAL_CAP * (int)(sizeof(VT) > 8 + sizeof(CT) ? sizeof(VT) : 8 + sizeof(CT));
  __shared__ __align__(16) unsigned char ub[UB];
  __shared__ uint32_t sT[RK_NT];
  __shared__ double sV[PM_MAXFAM * FAST_MAXENT];
  __shared__ double sR[PM_MAXGRPORD];
  __shared__ PT sPh[3 * PM_MAXORD];
  __shared__ uint32_t sPt[PRUNE ? PM_MAXPT : 1];  // least top block | mask of powers reaching it << 8
  __shared__ PowGen sPw[PRUNE ? PM_MAXPOW : 1];
  __shared__ uint8_t sPart[PM_MAXFAM * 32];
  __shared__ uint32_t wsum[BAL_TPB / 64];
  uint32_t *const tw = reinterpret_cast<uint32_t *>(ub);
  uint32_t *const tm = tw + BAL_CAP;
  CT *const tc = reinterpret_cast<CT *>(tm + BAL_CAP);
  VT *const contr = reinterpret_cast<VT *>(ub);
  const PermTab &T = *tab;
  const int tid = threadIdx.x;
  const Gen32 g0 = perm_gen(T, 0), g1 = perm_gen(T, 1), g2 = perm_gen(T, 2);
  const int n0 = T.gen[0].order, n1 = T.gen[1].order, n2 = T.gen[2].order;
  const uint32_t rcp0 = T.rcp0, rcp1 = T.rcp1;
  {
    for (int q = tid; q < RK_NT; q += BAL_TPB) sT[q] = A.rtab[q];
    if (tid < PM_MAXFAM * FAST_MAXENT) {
      const int f = tid / FAST_MAXENT, e = tid % FAST_MAXENT;
      sV[tid] = (f < T.nfam && e < T.fam[f].nent) ? T.fam[f].ent[e].val : 0.0;
    }
    if (tid < PM_MAXFAM * 32) {
      const int f = tid / 32;
      sPart[tid] = f < T.nfam ? T.fam[f].partner[tid % 32] : (uint8_t)0;
    }
    if (tid < PM_MAXGRPORD) sR[tid] = T.R[tid];
    if (tid < 3 * PM_MAXORD) {
      if constexpr (PH == 0) sPh[tid] = T.ph[tid].x;
      else sPh[tid] = T.ph[tid];
    }
    if constexpr (PRUNE) {
      const int np = 1 << T.pbits;
      for (int q = tid; q < np; q += BAL_TPB) {
        const uint2 e = ptab[q];
        sPt[q] = e.x | (e.y << 8);
      }
      if (tid < PM_MAXPOW) sPw[tid] = T.pow0[tid];
    }
  }
  const int psh = PRUNE ? A.L - T.pbits : 0;
  const uint32_t cb = bal_chunk(blockIdx.x, A.nchunk, A.xcdg);
  const uint32_t r0 = A.chunk[cb], nrow = A.chunk[cb + 1] - r0;
  // phase 1
  uint32_t s = 0, n = 0;
  const bool own = (uint32_t)tid < nrow;
  VT xi{};
  ST invR = 1;
  if (own) {
    xi = x[r0 + tid];
    invR = (ST)(1.0 / A.Rrow[r0 + tid]);
    s = A.keys[r0 + tid];
    const uint32_t ns = ~s & A.mask;
    for (int f = 0; f < T.nfam; f++) {
      const uint32_t r = perm_partner(T.fam[f], s), fm = T.fam[f].fmask;
      for (int e = 0; e < T.fam[f].nent; e++) n += __popc(ent_mask(T.fam[f].ent[e].locs, s, ns, r, A.mask) & fm);
    }
  }
  uint32_t total;
  // the scan's barrier also publishes the tables
  const uint32_t off = block_scan_u32(n, &total, wsum);
  if (own) {
    const uint32_t ns = ~s & A.mask;
    uint32_t o = off;
    for (int f = 0; f < T.nfam; f++) {
      const uint32_t r = perm_partner(T.fam[f], s), fm = T.fam[f].fmask;
      for (int e = 0; e < T.fam[f].nent; e++) {
        const int flip = T.fam[f].ent[e].flip;
        const uint32_t mb = (uint32_t)(f * FAST_MAXENT + e) | ((uint32_t)tid << 16);
        uint32_t bits = ent_mask(T.fam[f].ent[e].locs, s, ns, r, A.mask) & fm;
        while (bits) {
          const int p = __ffs(bits) - 1;
          bits &= bits - 1;
          uint32_t t = s;
          if (flip & 2) t ^= 1u << p;
          if (flip & 1) t ^= 1u << sPart[f * 32 + p];
          tw[o] = t;
          tm[o] = mb;
          o++;
        }
      }
    }
  }
  __syncthreads();
  // phase 2a: representative, coefficient and rank of each of the thread's
  // targets (its own slots only)
  auto finish = [&](uint32_t j, uint32_t rep, int be, int cnt) {
    const uint32_t m = tm[j];
    // exponents of element be = m0 + n0 (m1 + n1 m2) by reciprocal 24-bit
    // multiplies (be < 64; a runtime integer division costs quarter-rate
    // multiplies in a kernel bound by VALU issue)
    const uint32_t i1 = __umul24((uint32_t)be, rcp0) >> 16, i2 = __umul24(i1, rcp1) >> 16;
    const int m0 = be - (int)__umul24(i1, (uint32_t)n0), m1 = (int)(i1 - __umul24(i2, (uint32_t)n1)), m2 = (int)i2;
    const double R = sR[cnt - 1], v = sV[m & 63u];
    if constexpr (PH == 0) {
      tc[j] = (ST)(v * sPh[m0] * sPh[PM_MAXORD + m1] * sPh[2 * PM_MAXORD + m2] * R);
    } else {
      const double2 p = cmul(cmul(sPh[m0], sPh[PM_MAXORD + m1]), sPh[2 * PM_MAXORD + m2]);
      VT c;  // v conj(phase) R
      c.x = (ST)(v * p.x * R);
      c.y = (ST)(-v * p.y * R);
      tc[j] = c;
    }
    if (rank_domain(A, rep)) tw[j] = rank_full(sT, (rep & A.dmask) >> A.rlo, A.rng);
    else tm[j] = m | PMM_ABSENT;
  };
#pragma unroll 1
  for (int q = 0; q < BAL_K; q++) {
    const uint32_t j = tid + BAL_TPB * q;
    if (j < total) {
      uint32_t rep;
      int be, cnt;
      perm_canon<PRUNE, NPW>(A.L, g0, g1, g2, n0, n1, n2, sPt, psh, sPw, rcp0, tw[j], rep, be, cnt);
      finish(j, rep, be, cnt);
    }
  }
  // phase 2b (each thread reads back only its own slots)
  uint32_t rk[BAL_K], jj[BAL_K];
  bool ok[BAL_K];
  CT cf[BAL_K];
#pragma unroll
  for (int q = 0; q < BAL_K; q++) {
    const uint32_t j = tid + BAL_TPB * q;
    ok[q] = false;
    rk[q] = 0;
    cf[q] = CT{};
    if (j < total) {
      rk[q] = tw[j];
      cf[q] = tc[j];
      ok[q] = !(tm[j] & PMM_ABSENT);
    }
  }
  rank_lookup_many<BAL_K>(A, rk, ok, jj);
  VT xg[BAL_K];
#pragma unroll
  for (int q = 0; q < BAL_K; q++) xg[q] = x[jj[q] != HASH_EMPTY ? jj[q] : 0u];
  // every thread has read its targets before any contribution overlays them
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BAL_K; q++) {
    const uint32_t j = tid + BAL_TPB * q;
    const bool hit = jj[q] != HASH_EMPTY;
    VT cv;
    if constexpr (PH == 0) {
      cv.x = hit ? cf[q] * (ST)xg[q].x : (ST)0;
      cv.y = hit ? cf[q] * (ST)xg[q].y : (ST)0;
    } else {
      cv.x = hit ? cf[q].x * (ST)xg[q].x - cf[q].y * (ST)xg[q].y : (ST)0;
      cv.y = hit ? cf[q].x * (ST)xg[q].y + cf[q].y * (ST)xg[q].x : (ST)0;
    }
    if (j < total) contr[j] = cv;
  }
  __syncthreads();
  // phase 3
  if (own) {
    const uint32_t i = r0 + tid;
    const uint32_t ns = ~s & A.mask;
    ST dg = 0;
    for (int f = 0; f < T.nfam; f++) {
      const uint32_t r = perm_partner(T.fam[f], s), fm = T.fam[f].fmask;
      const uint32_t nr = ~r & A.mask;
      dg += (ST)(T.fam[f].diag[0] * __popc(ns & nr & fm) + T.fam[f].diag[1] * __popc(ns & r & fm) +
                 T.fam[f].diag[2] * __popc(s & nr & fm) + T.fam[f].diag[3] * __popc(s & r & fm));
    }
    ST orr = 0, oi = 0;
    for (uint32_t q = off; q < off + n; q++) {
      const VT cv = contr[q];
      orr += cv.x;
      oi += cv.y;
    }
    const ST yr = dg * (ST)xi.x + orr * invR;
    const ST yi = dg * (ST)xi.y + oi * invR;
    VT out;
    if (accumulate) {
      const VT y0 = y[i];
      out.x = y0.x + yr;
      out.y = y0.y + yi;
    } else {
      out.x = yr;
      out.y = yi;
    }
    y[i] = out;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// Two-site terms grouped into families (see the file comment); false when
// the operator does not have that form.
static bool perm_families(const edk_op *op, PermTab &tab) {
  const edk_basis *b = op->basis;
  const int L = b->L;
  if (op->nterms == 0) return false;
  // matrices compared by content: (row, column, value) of the 4 x 4 matrix
  typedef std::vector<std::pair<int, double2>> Content;
  std::vector<Content> mats;
  auto content = [&](int m) {
    Content c;
    const int *rp = &op->h_rowptr[op->h_mrow_off[m]];
    for (int loc = 0; loc < 4; loc++)
      for (int q = rp[loc]; q < rp[loc + 1]; q++) c.push_back({loc * 4 + op->h_col[q], op->h_val[q]});
    return c;
  };
  auto same = [](const Content &a, const Content &c) {
    if (a.size() != c.size()) return false;
    for (size_t i = 0; i < a.size(); i++)
      if (a[i].first != c[i].first || a[i].second.x != c[i].second.x || a[i].second.y != c[i].second.y)
        return false;
    return true;
  };
  std::vector<int> fam_mat;  // content id of each family
  tab.nfam = 0;
  for (int t = 0; t < op->nterms; t++) {
    const int *td = &op->h_tdesc[t * EDK_TDESC];
    if (td[0] != 2) return false;
    const int p1 = td[2], p2 = td[3];  // bit positions of the two sites (w == 1)
    if (p1 < 0 || p1 >= L || p2 < 0 || p2 >= L || p1 == p2) return false;
    const Content c = content(td[1]);
    int id = -1;
    for (size_t k = 0; k < mats.size() && id < 0; k++)
      if (same(mats[k], c)) id = (int)k;
    if (id < 0) {
      id = (int)mats.size();
      mats.push_back(c);
    }
    int f = -1;
    for (int k = 0; k < tab.nfam && f < 0; k++)
      if (fam_mat[k] == id && !(tab.fam[k].fmask & (1u << p1))) f = k;
    if (f < 0) {
      if (tab.nfam == PM_MAXFAM) return false;
      f = tab.nfam++;
      fam_mat.push_back(id);
      PermFam &F = tab.fam[f];
      memset(&F, 0, sizeof(F));
    }
    PermFam &F = tab.fam[f];
    F.fmask |= 1u << p1;
    F.partner[p1] = (uint8_t)p2;
    const int d = p1 - p2;  // bit p1 of r is bit p2 of s
    int g = -1;
    for (int k = 0; k < F.ng && g < 0; k++)
      if (F.gr[k] == (uint32_t)(d & 31)) g = k;
    if (g < 0) {
      if (F.ng == PM_MAXGRP) return false;
      g = F.ng++;
      F.gr[g] = (uint32_t)(d & 31);
      F.gm[g] = 0;
    }
    F.gm[g] |= 1u << p1;
  }
  for (int f = 0; f < tab.nfam; f++) {
    PermFam &F = tab.fam[f];
    F.nent = 0;
    for (const auto &e : mats[fam_mat[f]]) {
      const int loc = e.first / 4, c = e.first % 4;
      const double2 v = e.second;
      if (v.y != 0.0) return false;
      if (c == loc) {
        F.diag[loc] += v.x;
        continue;
      }
      int hit = -1;
      for (int q = 0; q < F.nent; q++)
        if (F.ent[q].flip == (loc ^ c) && F.ent[q].val == v.x) hit = q;
      if (hit >= 0) {
        F.ent[hit].locs |= 1 << loc;
        continue;
      }
      if (F.nent >= FAST_MAXENT) return false;
      F.ent[F.nent].locs = 1 << loc;
      F.ent[F.nent].flip = loc ^ c;
      F.ent[F.nent].val = v.x;
      F.nent++;
    }
  }
  return true;
}

static uint32_t rotl32_host(uint32_t v, uint32_t k) { return (v << k) | (v >> ((32u - k) & 31u)); }

static uint32_t perm_partner_host(const PermFam &F, uint32_t s) {
  uint32_t r = 0;
  for (int g = 0; g < F.ng; g++) r |= rotl32_host(s, F.gr[g]) & F.gm[g];
  return r;
}

static uint32_t perm_apply_host(int L, const PermGen &G, uint32_t s) {
  if (G.rev) s = __builtin_bitreverse32(s) >> (32 - L);
  uint32_t o = 0;
  for (int q = 0; q < G.n; q++) o |= rotl32_host(s & G.m[q], G.rl[q]);
  return o ^ G.inv;
}

// Generator 0's powers as PowGen (bit permutation by rotation groups, then
// the accumulated inversion g^a(0)); false when a power needs more than
// EDK_MAXSHIFTS groups (then the walk is not pruned).
static bool perm_powers(int L, PermTab &tab, int &npw) {
  const PermGen &G = tab.gen[0];
  if (G.order > PM_MAXPOW) return false;
  uint32_t img[32];  // g^a of bit p, without the inversion
  for (int p = 0; p < L; p++) img[p] = 1u << p;
  uint32_t inv = 0;
  for (int a = 0; a < G.order; a++) {
    PowGen &P = tab.pow0[a];
    memset(&P, 0, sizeof(P));
    P.inv = inv;
    int ng = 0;
    for (int p = 0; p < L; p++) {
      if (__builtin_popcount(img[p]) != 1) return false;
      const uint32_t r = (uint32_t)(__builtin_ctz(img[p]) - p) & 31u;
      int g = -1;
      for (int k = 0; k < ng && g < 0; k++)
        if (P.rl[k] == r) g = k;
      if (g < 0) {
        if (ng == EDK_MAXSHIFTS) return false;
        g = ng++;
        P.rl[g] = r;
      }
      P.m[g] |= 1u << p;
    }
    // next power: apply g to every bit image (the inversion is tracked apart)
    for (int p = 0; p < L; p++) img[p] = perm_apply_host(L, G, img[p]) ^ G.inv;
    inv = perm_apply_host(L, G, inv);
  }
  npw = 1;
  for (int a = 0; a < G.order; a++)
    for (int q = 0; q < EDK_MAXSHIFTS; q++)
      if (tab.pow0[a].m[q]) npw = std::max(npw, q + 1);
  tab.rcp0 = (65536u + (uint32_t)G.order - 1) / (uint32_t)G.order;
  for (uint32_t e = 0; e < PM_MAXGRPORD; e++)
    if (((e * tab.rcp0) >> 16) != e / (uint32_t)G.order) return false;
  return true;
}

// The basis' generators as PermGen; false for a form this kernel lacks.
static bool perm_gens(const edk_basis *b, PermTab &tab) {
  const int L = b->L;
  const uint32_t mask = L == 32 ? 0xffffffffu : ((1u << L) - 1);
  tab.ngen = b->ngen;
  for (int g = 0; g < 3; g++) {
    PermGen &G = tab.gen[g];
    memset(&G, 0, sizeof(G));
    G.n = 1;
    G.m[0] = mask;  // unused generators: the identity, order 1
    G.order = 1;
    if (g >= b->ngen) continue;
    const int *d = &b->h_gdesc[EDK_GDESC * g];
    G.order = d[1];
    G.inv = (uint32_t)b->h_gdata[d[3]];
    switch (d[0]) {
      case GEN_IDENT: break;
      case GEN_REV: G.rev = 1; break;
      case GEN_ROTL: {
        const int k = d[2];
        if (k > 0 && k < L) {
          G.n = 2;
          G.m[0] = mask >> k;  // bits that move up by k
          G.rl[0] = (uint32_t)k;
          G.m[1] = mask & ~(mask >> k);  // bits that wrap down by L - k
          G.rl[1] = (uint32_t)((k - L) & 31);
        }
        break;
      }
      case GEN_SHIFTS: {
        G.n = d[5];
        if (G.n < 1 || G.n > EDK_MAXSHIFTS) return false;
        for (int q = 0; q < G.n; q++) {
          G.m[q] = (uint32_t)b->h_gdata[d[4] + q];
          G.rl[q] = (uint32_t)(b->h_gshift[d[4] + q] & 31);
        }
        break;
      }
      default: return false;
    }
  }
  return true;
}

static void perm_free(PermState *P) {
  (void)hipFree(P->d_tab);
  (void)hipFree(P->d_keys32);
  (void)hipFree(P->d_chunk);
  (void)hipFree(P->d_gran);
  (void)hipFree(P->d_rtab);
}

static bool perm_build(edk_op *op, PermState *P) {
  const edk_basis *b = op->basis;
  if (b->kind != EDK_ABELIAN || b->base != 2 || b->w != 1 || b->L > 32 || b->uploaded) return false;
  if (b->D == 0 || b->D >= (1ll << 32) - 1) return false;
  if (b->ngen < 1 || b->ngen > 3 || b->order > PM_MAXGRPORD) return false;
  for (int g = 0; g < b->ngen; g++) {
    const int type = b->h_gdesc[EDK_GDESC * g], ord = b->h_gdesc[EDK_GDESC * g + 1];
    if (type == GEN_BENES || ord > PM_MAXORD || ord < 1) return false;
  }
  PermTab tab;
  memset(&tab, 0, sizeof(tab));
  if (!perm_families(op, tab) || !perm_gens(b, tab)) return false;
  // reciprocals that split an element number (< 64) into the exponents of
  // generators 0 and 1, checked exhaustively
  tab.rcp0 = (65536u + (uint32_t)tab.gen[0].order - 1) / (uint32_t)tab.gen[0].order;
  tab.rcp1 = (65536u + (uint32_t)tab.gen[1].order - 1) / (uint32_t)tab.gen[1].order;
  for (uint32_t e = 0; e < PM_MAXGRPORD; e++)
    if (((e * tab.rcp0) >> 16) != e / (uint32_t)tab.gen[0].order ||
        ((e * tab.rcp1) >> 16) != e / (uint32_t)tab.gen[1].order)
      return false;
  // the pruned walk when the basis has its table and it fits in LDS
  P->ptab = (b->d_ptab && b->pbits >= 1 && (1 << b->pbits) <= PM_MAXPT && perm_powers(b->L, tab, P->npw)) ? b->d_ptab
                                                                                                : nullptr;
  tab.pbits = P->ptab ? b->pbits : 0;
  for (int i = 0; i < b->order && i < PM_MAXGRPORD; i++) tab.R[i] = b->h_Rtab[i];
  P->real = true;
  for (int g = 0; g < 3; g++)
    for (int m = 0; m < PM_MAXORD; m++) {
      double2 p = make_double2(1.0, 0.0);
      if (g < b->ngen && m < tab.gen[g].order) p = b->h_ph[b->h_gdesc[EDK_GDESC * g + 6] + m];
      tab.ph[g * PM_MAXORD + m] = p;
      P->real = P->real && p.y == 0.0;
    }
  edk_ctx *ctx = op->ctx;
  const int64_t D = b->D;
  std::vector<uint64_t> k64(D);
  if (hipMemcpyAsync(k64.data(), b->d_keys, D * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    return false;
  std::vector<uint32_t> keys(D);
  for (int64_t i = 0; i < D; i++) keys[i] = (uint32_t)k64[i];
  k64.clear();
  k64.shrink_to_fit();
  FastArgs &A = P->A;
  memset(&A, 0, sizeof(A));
  A.kind = EDK_ABELIAN;
  A.L = b->L;
  A.A = 1;
  A.len = b->L;
  A.mask = b->L == 32 ? 0xffffffffu : ((1u << b->L) - 1);
  A.D = (uint32_t)D;
  A.Rrow = b->d_R;
  A.xcdg = 512;
  if (hipMalloc((void **)&P->d_keys32, D * sizeof(uint32_t)) != hipSuccess ||
      hipMemcpyAsync(P->d_keys32, keys.data(), D * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream) !=
          hipSuccess)
    return false;
  A.keys = P->d_keys32;
  if (!edk_rank_build(ctx, b->L, keys, P->d_keys32, A, &P->d_gran, &P->d_rtab) || !A.gran) return false;
  // balanced chunks: whole rows, at most BAL_CAP targets and BAL_TPB rows
  std::vector<uint32_t> cs;
  cs.push_back(0);
  uint32_t acc = 0, rows = 0;
  for (int64_t i = 0; i < D; i++) {
    const uint32_t s = keys[i], ns = ~s & A.mask;
    uint32_t c = 0;
    for (int f = 0; f < tab.nfam; f++) {
      const PermFam &F = tab.fam[f];
      const uint32_t r = perm_partner_host(F, s), nr = ~r & A.mask;
      for (int e = 0; e < F.nent; e++) {
        const int l = F.ent[e].locs;
        uint32_t m = 0;
        if (l & 1) m |= ns & nr;
        if (l & 2) m |= ns & r;
        if (l & 4) m |= s & nr;
        if (l & 8) m |= s & r;
        c += (uint32_t)__builtin_popcount(m & F.fmask);
      }
    }
    if (c > BAL_CAP) return false;
    if (rows == BAL_TPB || acc + c > BAL_CAP) {
      cs.push_back((uint32_t)i);
      acc = 0;
      rows = 0;
    }
    acc += c;
    rows++;
  }
  cs.push_back((uint32_t)D);
  if (hipMalloc((void **)&P->d_chunk, cs.size() * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void **)&P->d_tab, sizeof(PermTab)) != hipSuccess ||
      hipMemcpyAsync(P->d_chunk, cs.data(), cs.size() * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream) !=
          hipSuccess ||
      hipMemcpyAsync(P->d_tab, &tab, sizeof(PermTab), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    return false;
  A.chunk = P->d_chunk;
  A.nchunk = (uint32_t)cs.size() - 1;
  return true;
}

static std::mutex g_perm_mu;

// One-time setup, published by the single store of pstate (launches read it
// without the lock, as for the ring kernels).
bool edk_perm_prepare(edk_op *op) {
  const int st = op->pstate.load(std::memory_order_acquire);
  if (st) return st > 0;
  std::lock_guard<std::mutex> lk(g_perm_mu);
  if (op->pstate.load(std::memory_order_relaxed)) return op->pstate.load(std::memory_order_relaxed) > 0;
  PermState *P = new PermState();
  const bool ok = perm_build(op, P);
  if (ok) {
    op->perm = P;
  } else {
    perm_free(P);
    delete P;
    (void)hipGetLastError();
  }
  op->pstate.store(ok ? 1 : -1, std::memory_order_release);
  return ok;
}

void edk_perm_release(edk_op *op) {
  PermState *P = (PermState *)op->perm;
  if (!P) return;
  perm_free(P);
  delete P;
  op->perm = nullptr;
  op->pstate.store(0, std::memory_order_release);
}

template <int PH, bool PRUNE, int NPW>
static void launch_perm(const PermState &P, int dtype, void *y, const void *x, int acc, hipStream_t st) {
  if (dtype == EDK_C64)
    k_matvec_perm<PH, PRUNE, NPW, double2, double>
        <<<P.A.nchunk, BAL_TPB, 0, st>>>(P.A, P.d_tab, P.ptab, (const double2 *)x, (double2 *)y, acc);
  else
    k_matvec_perm<PH, PRUNE, NPW, float2, float>
        <<<P.A.nchunk, BAL_TPB, 0, st>>>(P.A, P.d_tab, P.ptab, (const float2 *)x, (float2 *)y, acc);
}

bool edk_matvec_perm(edk_op *op, int dtype, void *y, const void *x, int accumulate, hipStream_t st) {
  if (op->disable_fast || op->basis->kind != EDK_ABELIAN) return false;
  if (op->fast) {
    // a schedule forced or tuned for one vector: only the tile form is this kernel
    const int v = ((const FastState *)op->fast)->tuned[dtype == EDK_C64 ? 0 : 1][1];
    if (v != EDK_VARIANT_AUTO && v != EDK_VARIANT_TILE) return false;
  }
  if (!edk_perm_prepare(op)) return false;
  const PermState &P = *(const PermState *)op->perm;
  if (!P.ptab) {
    if (P.real) launch_perm<0, false, 2>(P, dtype, y, x, accumulate, st);
    else launch_perm<2, false, 2>(P, dtype, y, x, accumulate, st);
  } else if (P.npw <= 2) {
    if (P.real) launch_perm<0, true, 2>(P, dtype, y, x, accumulate, st);
    else launch_perm<2, true, 2>(P, dtype, y, x, accumulate, st);
  } else {
    if (P.real) launch_perm<0, true, EDK_MAXSHIFTS>(P, dtype, y, x, accumulate, st);
    else launch_perm<2, true, EDK_MAXSHIFTS>(P, dtype, y, x, accumulate, st);
  }
  return true;
}
