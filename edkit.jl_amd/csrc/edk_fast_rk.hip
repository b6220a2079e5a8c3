// edk_fast_rk.hip -- one-vector gather kernel over the rank directory.
//
// The balanced chunk schedule of k_matvec_bal (one block per chunk of whole
// rows holding at most 1024 targets) with the run-length canonicalisers and
// real phases, where positions come from the rank directory (edk_fast.h)
// instead of the key table:
//   phase 1  row threads write their targets (word and entry to LDS arrays)
//            and their own row's rank;
//   phase 2a every thread canonicalises four targets.  A target that is its
//            own representative and differs from its row by an exchange
//            inside a window of at most three domain bits takes the row's
//            rank with the window's colex terms exchanged; the others are
//            queued: words with more than two longest zero runs for the full
//            canonicaliser, the rest for a full rank;
//   phase 2b the queues, spread over the whole block;
//   phase 2c granule loads, then x gathers, all four targets in flight; the
//            contributions then take the place of the target arrays;
//   phase 3  row threads sum their contributions in generator order, add the
//            diagonal and write y.
#include "edk_fast.h"

// target meta word: entry (bits 0-5), first site's bit (6-10), family
// distance (11-12), exchange rule (13), absent (14), row in the chunk (16-23),
// the norm-table index of the representative's factor (24-29) and its sign (30)
#define RKM_XR 0x2000u
#define RKM_ABSENT 0x4000u
#define RKM_NEG 0x40000000u

template <int KIND, int PH, int LT, typename VT, typename ST>
// 8 waves per SIMD (64 VGPRs); the real-sign L = 32 form spills there and keeps 6
__global__ void __launch_bounds__(BAL_TPB, (PH == 1 && LT == 32) ? 6 : 8)
    k_matvec_rk(FastArgs A, const VT *__restrict__ x, VT *__restrict__ y, int accumulate) {
  static_assert(PH != 2, "real phases only");
  static_assert(LT == 32 || LT == 1, "run-length canonicalisers only");
  // per target: word (then representative, then rank) and meta word, as two
  // arrays (conflict-free 4-byte accesses).  A real coefficient is the
  // entry's value times +-R from the norm table, so the meta word carries the
  // table index and sign instead of an 8-byte coefficient.  After the last
  // reads of phase 2c the contributions overlay the arrays, in NH passes of
  // HCAP targets (ComplexF64: two passes of 512).  The arrays stay at 8 KB:
  // 18 KB of LDS for the workgroup, 8 workgroups per CU.
  constexpr int NH = sizeof(VT) > 8 ? 2 : 1;
  constexpr int HCAP = BAL_CAP / NH;
  static_assert(HCAP == BAL_TPB * (BAL_K / NH) && HCAP * (int)sizeof(VT) <= BAL_CAP * 8,
                "a pass holds whole target slots and fits the target arrays");
  // orbit lengths are at most 32 (32-bit keys): norms for both parities
  __shared__ double sR[2 * 32];
  __shared__ double sV[FAST_MAXFAM * FAST_MAXENT];
  __shared__ uint32_t sT[RK_NT];
  __shared__ __align__(16) uint32_t ub[2 * BAL_CAP];
  uint32_t *const tw = ub;
  uint32_t *const tm = ub + BAL_CAP;
  VT *const contr = reinterpret_cast<VT *>(ub);
  __shared__ uint16_t rq[BAL_CAP];  // brute-force queue from the front, rank queue from the back
  __shared__ uint32_t rrow[BAL_TPB];
  __shared__ uint32_t wsum[BAL_TPB / 64];
  __shared__ uint32_t nbq, nrq;
  const FastTab &T = *A.tab;
  const int tid = threadIdx.x;
  if (tid == 0) {
    nbq = 0;
    nrq = 0;
  }
  {
    const int nr = (KIND == EDK_TRANSLATION_PARITY ? 2 : 1) * A.len;
    if (tid < nr) sR[tid] = A.Rtab[tid];
    if (tid < FAST_MAXFAM * FAST_MAXENT) {
      const int f = tid / FAST_MAXENT, e = tid % FAST_MAXENT;
      sV[tid] = (f < T.nfam && e < T.fam[f].nent) ? T.fam[f].ent[e].val : 0.0;
    }
    for (int q = tid; q < RK_NT; q += BAL_TPB) sT[q] = A.rtab[q];
  }
  FastArgs B = A;
  B.Rtab = sR;
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
      const uint32_t r = rotr_lt<LT>(s, A.L - T.fam[f].d, A.L, A.mask);
      for (int e = 0; e < T.fam[f].nent; e++) n += __popc(ent_mask(T.fam[f].ent[e].locs, s, ns, r, A.mask));
    }
  }
  uint32_t total;
  // the scan's barrier also publishes the tables
  const uint32_t off = block_scan_u32(n, &total, wsum);
  if (own) {
    rrow[tid] = rank_full(sT, (s & A.dmask) >> A.rlo, A.rng);
    const uint32_t ns = ~s & A.mask;
    uint32_t o = off;
    for (int f = 0; f < T.nfam; f++) {
      const int d = T.fam[f].d;
      const uint32_t r = rotr_lt<LT>(s, A.L - d, A.L, A.mask);
      for (int e = 0; e < T.fam[f].nent; e++) {
        const int flip = T.fam[f].ent[e].flip;
        // the exchange rule: both sites flip, antiparallel pairs only (the
        // popcount is kept), a window of at most three bits
        const uint32_t xr = (flip == 3 && (T.fam[f].ent[e].locs & 9) == 0 && d <= 2) ? RKM_XR : 0u;
        const uint32_t mb = (uint32_t)(f * FAST_MAXENT + e) | ((uint32_t)(d & 3) << 11) | xr | ((uint32_t)tid << 16);
        uint32_t bits = ent_mask(T.fam[f].ent[e].locs, s, ns, r, A.mask);
        while (bits) {
          const int p = __ffs(bits) - 1;
          bits &= bits - 1;
          int p2 = p - d;
          if (LT == 32) p2 &= 31;
          else if (p2 < 0) p2 += A.L;
          uint32_t t = s;
          if (flip & 2) t ^= 1u << p;
          if (flip & 1) t ^= 1u << p2;
          tw[o] = t;
          tm[o] = mb | ((uint32_t)p << 6);
          o++;
        }
      }
    }
  }
  __syncthreads();
  // phase 2a
#pragma unroll 1
  for (int q = 0; q < BAL_K; q++) {
    const uint32_t j = tid + BAL_TPB * q;
    if (j < total) {
      const uint32_t t = tw[j], m = tm[j];
      uint32_t r = 0;
      double cr = 0.0, ci = 0.0;
      int ri = 0;
      const int st = canon_runs<KIND, PH, LT>(B, t, r, cr, ci, &ri);
      if (st > 0) {
        tm[j] = m | ((uint32_t)ri << 24) | (cr < 0.0 ? RKM_NEG : 0u);
        const int p = (m >> 6) & 31, d = (m >> 11) & 3, lo2 = p - d;
        if ((m & RKM_XR) && r == t && lo2 >= A.rlo && p < A.rlo + A.rbits) {
          const uint32_t sw = t ^ ((1u << p) | (1u << lo2));  // the row's word
          const int a = lo2 - A.rlo;
          const int c = __popc((sw >> A.rlo) & ((1u << a) - 1u));
          const uint32_t rs = rrow[m >> 16];
          uint32_t rk;
          if (d == 1) {
            // the set bit moves from a to a + 1 (or back): C(a + 1, c + 1) - C(a, c + 1) = C(a, c)
            const uint32_t cb_ = rank_binom(sT, a, c);
            rk = ((t >> p) & 1u) ? rs + cb_ : rs - cb_;
          } else {
            const uint32_t wm = (2u << d) - 1u;
            rk = rs - rank_window(sT, (sw >> lo2) & wm, a, c) + rank_window(sT, (t >> lo2) & wm, a, c);
          }
          tw[j] = rk;
        } else {
          tw[j] = r;
          rq[BAL_CAP - 1 - atomicAdd(&nrq, 1u)] = (uint16_t)j;
        }
      } else if (st == 0) {
        tm[j] = m | RKM_ABSENT;
      } else {
        rq[atomicAdd(&nbq, 1u)] = (uint16_t)j;
      }
    }
  }
  __syncthreads();
  // phase 2b: brute-force canonicalisation (then a full rank), then full ranks
  {
    const uint32_t nb = nbq, ne = nbq + nrq;
    for (uint32_t e = tid; e < ne; e += BAL_TPB) {
      const bool brute = e < nb;
      const uint32_t j = brute ? rq[e] : rq[BAL_CAP - 1 - (e - nb)];
      uint32_t r = tw[j];
      bool okb = true;
      if (brute) {
        double cr = 0.0, ci = 0.0;
        int ri = 0;
        const uint32_t t = r;
        okb = (LT == 32) ? fast_canon32<KIND, PH>(B, t, r, cr, ci, &ri)
                         : fast_canon_w<KIND, PH>(B, t, r, cr, ci, &ri);
        if (okb) tm[j] = (tm[j] & 0x00ffffffu) | ((uint32_t)ri << 24) | (cr < 0.0 ? RKM_NEG : 0u);
      }
      okb = okb && rank_domain(A, r);
      if (okb) tw[j] = rank_full(sT, (r & A.dmask) >> A.rlo, A.rng);
      else tm[j] |= RKM_ABSENT;
    }
  }
  __syncthreads();
  // phase 2c
  uint32_t rk[BAL_K], jj[BAL_K];
  bool ok[BAL_K];
  ST cf[BAL_K];
#pragma unroll
  for (int q = 0; q < BAL_K; q++) {
    const uint32_t j = tid + BAL_TPB * q;
    ok[q] = false;
    rk[q] = 0;
    cf[q] = 0;
    if (j < total) {
      rk[q] = tw[j];
      const uint32_t m = tm[j];
      ok[q] = !(m & RKM_ABSENT);
      // the coefficient as phase 2 formed it: value * (sign * R)
      const double R = sR[(m >> 24) & 63u];
      cf[q] = (ST)(sV[m & 63u] * ((m & RKM_NEG) ? -R : R));
    }
  }
  rank_lookup_many<BAL_K>(A, rk, ok, jj);
  VT xg[BAL_K];
#pragma unroll
  for (int q = 0; q < BAL_K; q++) xg[q] = x[jj[q] != HASH_EMPTY ? jj[q] : 0u];
  // every thread has read its targets before any contribution overlays them
  __syncthreads();
  // phase 3: the diagonal, then the row's contributions in generator order,
  // one pass of HCAP target slots at a time (the sums run on across passes)
  const uint32_t qe = off + n;
  uint32_t qq = off;
  ST dg = 0, orr = 0, oi = 0;
  if (own) {
    const uint32_t ns = ~s & A.mask;
    for (int f = 0; f < T.nfam; f++) {
      const uint32_t r = rotr_lt<LT>(s, A.L - T.fam[f].d, A.L, A.mask);
      const uint32_t nr = ~r & A.mask;
      dg += (ST)(T.fam[f].diag[0] * __popc(ns & nr) + T.fam[f].diag[1] * __popc(ns & r) +
                 T.fam[f].diag[2] * __popc(s & nr) + T.fam[f].diag[3] * __popc(s & r));
    }
  }
#pragma unroll
  for (int h = 0; h < NH; h++) {
#pragma unroll
    for (int q = h * (BAL_K / NH); q < (h + 1) * (BAL_K / NH); q++) {
      const uint32_t j = tid + BAL_TPB * q;
      const bool hit = jj[q] != HASH_EMPTY;
      VT cv;
      cv.x = hit ? cf[q] * (ST)xg[q].x : (ST)0;
      cv.y = hit ? cf[q] * (ST)xg[q].y : (ST)0;
      if (j < total) contr[j - h * HCAP] = cv;
    }
    __syncthreads();
    if (own) {
      const uint32_t he = min(qe, (uint32_t)((h + 1) * HCAP));
      const uint32_t hb = (uint32_t)(h * HCAP);
      for (; qq + 4 <= he; qq += 4) {
        const VT c0 = contr[qq - hb], c1 = contr[qq + 1 - hb], c2 = contr[qq + 2 - hb], c3 = contr[qq + 3 - hb];
        orr += c0.x;
        oi += c0.y;
        orr += c1.x;
        oi += c1.y;
        orr += c2.x;
        oi += c2.y;
        orr += c3.x;
        oi += c3.y;
      }
      for (; qq < he; qq++) {
        const VT cv = contr[qq - hb];
        orr += cv.x;
        oi += cv.y;
      }
    }
    // the next pass overwrites the contributions
    if (h + 1 < NH) __syncthreads();
  }
  if (own) {
    const uint32_t i = r0 + tid;
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

template <int KIND, int PH, int LT>
static void launch_rk(const FastArgs &A, int dtype, void *y, const void *x, int acc, hipStream_t st) {
  if (dtype == EDK_C64)
    k_matvec_rk<KIND, PH, LT, double2, double><<<A.nchunk, BAL_TPB, 0, st>>>(A, (const double2 *)x, (double2 *)y, acc);
  else
    k_matvec_rk<KIND, PH, LT, float2, float><<<A.nchunk, BAL_TPB, 0, st>>>(A, (const float2 *)x, (float2 *)y, acc);
}

template <int KIND, int PH>
static bool launch_rk_lt(const FastArgs &A, int dtype, void *y, const void *x, int acc, hipStream_t st) {
  if (A.L == 32 && A.A == 1) launch_rk<KIND, PH, 32>(A, dtype, y, x, acc, st);
  else if (A.A == 1 && A.L >= 16) launch_rk<KIND, PH, 1>(A, dtype, y, x, acc, st);
  else return false;
  return true;
}

bool edk_fast_launch_rk(const FastArgs &A, int dtype, void *y, const void *x, int acc, hipStream_t st) {
  if (!A.gran || !A.chunk) return false;
  switch (A.kind) {
    case EDK_TRANSLATION_PARITY:
      if (A.kp == 1 && A.P == 1) return launch_rk_lt<EDK_TRANSLATION_PARITY, 0>(A, dtype, y, x, acc, st);
      return launch_rk_lt<EDK_TRANSLATION_PARITY, 1>(A, dtype, y, x, acc, st);
    case EDK_TRANSLATIONAL:
      if (A.kp == 0) return launch_rk_lt<EDK_TRANSLATIONAL, 0>(A, dtype, y, x, acc, st);
      if (2 * A.kp == A.len) return launch_rk_lt<EDK_TRANSLATIONAL, 1>(A, dtype, y, x, acc, st);
      return false;
    default: return false;
  }
}
