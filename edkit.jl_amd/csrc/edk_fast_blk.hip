// edk_fast_blk.hip -- specialised gather kernel for blocks of NV vectors
// (A10: mul!(Y, H, X), Operator.jl:322-328, 382-394).
#include "edk_fast.h"

// Block of NV vectors, row-interleaved (element (i, c) at x[i*NV + c]): the
// same chunks and phases as k_matvec_bal, but phase 2 keeps each target's
// (position, coefficient) in LDS, and phase 3 gives every row consecutive
// threads of one or two columns each, so each representative lookup serves
// all NV columns and the NV gathered values of a target form one contiguous
// segment.
template <int KIND, int PH, int LT, typename VT, typename ST, int NV>
__device__ __forceinline__ void bal_blk_body(const FastArgs &A, uint32_t cb, const VT *__restrict__ x,
                                             VT *__restrict__ y, int accumulate) {
  constexpr bool CPLX = (PH == 2);
  // orbit lengths are at most 32 (32-bit keys): norms for both parities and
  // one phase per rotation count
  __shared__ double sR[2 * 32];
  __shared__ double2 sP[PH != 0 ? 32 : 1];
  __shared__ double sV[FAST_MAXFAM * FAST_MAXENT];
  __shared__ uint32_t tgt[BAL_CAP];
  __shared__ uint8_t ent[BAL_CAP];
  __shared__ ST tcr[BAL_CAP], tci[CPLX ? BAL_CAP : 1];
  __shared__ uint32_t roff[BAL_TPB + 1];  // row r's targets: [roff[r], roff[r + 1])
  __shared__ ST rdg[BAL_TPB], rinv[BAL_TPB];
  __shared__ uint32_t wsum[BAL_TPB / 64];
  // run-length canonicaliser (unit cell 1, 16 <= L <= 32) with a block-wide
  // full search for the targets it defers
  constexpr bool RUNS = (LT == 32 || LT == 1) && KIND != EDK_PROJECTED;
  __shared__ uint16_t dq[RUNS ? BAL_CAP : 1];
  __shared__ uint32_t ndef;
  const FastTab &T = *A.tab;
  const int tid = threadIdx.x;
  if (RUNS && tid == 0) ndef = 0;
  if (KIND != EDK_PROJECTED) {
    int nr = (KIND == EDK_TRANSLATION_PARITY ? 2 : 1) * A.len;
    if (tid < nr) sR[tid] = A.Rtab[tid];
    if (PH != 0 && KIND == EDK_TRANSLATIONAL && tid < A.len) sP[tid] = A.ph[tid];
  }
  if (tid < FAST_MAXFAM * FAST_MAXENT) {
    int f = tid / FAST_MAXENT, e = tid % FAST_MAXENT;
    sV[tid] = (f < T.nfam && e < T.fam[f].nent) ? T.fam[f].ent[e].val : 0.0;
  }
  FastArgs B = A;
  B.Rtab = sR;
  if (PH != 0 && KIND == EDK_TRANSLATIONAL) B.ph = sP;
  const uint32_t r0 = A.chunk[cb], nrow = A.chunk[cb + 1] - r0;
  // phase 1: row threads count and write their targets; diagonal and 1/R_i
  uint32_t s = 0, n = 0;
  const bool own = (uint32_t)tid < nrow;
  if (own) {
    s = A.keys[r0 + tid];
    const uint32_t ns = ~s & A.mask;
    for (int f = 0; f < T.nfam; f++) {
      uint32_t r = rotr_lt<LT>(s, A.L - T.fam[f].d, A.L, A.mask);
      for (int e = 0; e < T.fam[f].nent; e++) n += __popc(ent_mask(T.fam[f].ent[e].locs, s, ns, r, A.mask));
    }
  }
  uint32_t total;
  const uint32_t off = block_scan_u32(n, &total, wsum);
  if (own) {
    const uint32_t ns = ~s & A.mask;
    uint32_t o = off;
    ST dg = 0;
    for (int f = 0; f < T.nfam; f++) {
      const int d = T.fam[f].d;
      uint32_t r = rotr_lt<LT>(s, A.L - d, A.L, A.mask);
      uint32_t nr = ~r & A.mask;
      dg += (ST)(T.fam[f].diag[0] * __popc(ns & nr) + T.fam[f].diag[1] * __popc(ns & r) +
                 T.fam[f].diag[2] * __popc(s & nr) + T.fam[f].diag[3] * __popc(s & r));
      for (int e = 0; e < T.fam[f].nent; e++) {
        const int flip = T.fam[f].ent[e].flip;
        uint32_t bits = ent_mask(T.fam[f].ent[e].locs, s, ns, r, A.mask);
        while (bits) {
          int p = __ffs(bits) - 1;
          bits &= bits - 1;
          int p2 = p - d;
          if (LT == 32) p2 &= 31;
          else if (p2 < 0) p2 += A.L;
          uint32_t t = s;
          if (flip & 2) t ^= 1u << p;
          if (flip & 1) t ^= 1u << p2;
          tgt[o] = t;
          ent[o] = (uint8_t)(f * FAST_MAXENT + e);
          o++;
        }
      }
    }
    roff[tid] = off;
    if ((uint32_t)tid == nrow - 1) roff[nrow] = total;
    rdg[tid] = dg;
    rinv[tid] = (KIND == EDK_PROJECTED) ? (ST)1 : (ST)(1.0 / A.Rrow[r0 + tid]);
  }
  __syncthreads();
  // phase 2: BAL_K targets per thread -> (position, coefficient)
  uint32_t rep[BAL_K], jj[BAL_K];
  bool ok[BAL_K];
  ST cfr[BAL_K], cfi[CPLX ? BAL_K : 1];
  if constexpr (RUNS) {
    // one target at a time, results to LDS; ent = 0xff marks a representative
    // outside the sector
#pragma unroll 1
    for (int q = 0; q < BAL_K; q++) {
      const uint32_t j = tid + BAL_TPB * q;
      if (j < total) {
        uint32_t r;
        double cr = 0.0, ci = 0.0;
        const int st = canon_runs<KIND, PH, LT>(B, tgt[j], r, cr, ci);
        if (st > 0) {
          tgt[j] = r;
          const double v = sV[ent[j]];
          tcr[j] = (ST)(v * cr);
          if (CPLX) tci[j] = (ST)(v * ci);
        } else if (st == 0) {
          ent[j] = 0xff;
        } else {
          dq[atomicAdd(&ndef, 1u)] = (uint16_t)j;
        }
      }
    }
    __syncthreads();
    const uint32_t nd = ndef;
    for (uint32_t d = tid; d < nd; d += BAL_TPB) {
      const uint32_t j = dq[d];
      uint32_t r = 0;
      double cr = 0.0, ci = 0.0;
      const bool okb = (LT == 32) ? fast_canon32<KIND, PH>(B, tgt[j], r, cr, ci)
                                  : fast_canon_w<KIND, PH>(B, tgt[j], r, cr, ci);
      if (okb) {
        tgt[j] = r;
        const double v = sV[ent[j]];
        tcr[j] = (ST)(v * cr);
        if (CPLX) tci[j] = (ST)(v * ci);
      } else {
        ent[j] = 0xff;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BAL_K; q++) {
      const uint32_t j = tid + BAL_TPB * q;
      ok[q] = j < total && ent[j] != 0xff;
      rep[q] = ok[q] ? tgt[j] : 0u;
      cfr[q] = ok[q] ? tcr[j] : (ST)0;
      if (CPLX) cfi[q] = ok[q] ? tci[j] : (ST)0;
    }
  } else {
#pragma unroll
    for (int q = 0; q < BAL_K; q++) {
      const uint32_t j = tid + BAL_TPB * q;
      ok[q] = j < total;
      uint32_t t = ok[q] ? tgt[j] : 0u;
      double cr = 0.0, ci = 0.0;
      rep[q] = 0;
      if (ok[q]) {
        if (LT == 32) ok[q] = fast_canon32<KIND, PH>(B, t, rep[q], cr, ci);
        else ok[q] = fast_canon<KIND, PH>(B, t, rep[q], cr, ci);
      }
      double v = ok[q] ? sV[ent[j]] : 0.0;
      cfr[q] = (ST)(v * cr);
      if (CPLX) cfi[q] = (ST)(v * ci);
    }
  }
  region_lookup_many<BAL_K>(A, rep, ok, jj);
#pragma unroll
  for (int q = 0; q < BAL_K; q++) {
    const uint32_t j = tid + BAL_TPB * q;
    if (j < total) {
      tgt[j] = jj[q];
      tcr[j] = cfr[q];
      if (CPLX) tci[j] = cfi[q];
    }
  }
  __syncthreads();
  // phase 3: NV / CPL threads per row, each summing CPL columns' targets in
  // generator order.  A block of 8 takes two columns a thread, so a chunk's
  // row-column pairs fit one pass of the block (with one column a thread, a
  // chunk of more than 32 rows needs a second pass that the first wave runs
  // alone, lengthening the block's chain of gathers)
  constexpr int CPL = NV == 8 ? 2 : 1;
  constexpr int LPR = NV / CPL;
  constexpr int GF = (CPL == 2 && sizeof(VT) > 8) ? 4 : 8;  // targets' gathers in flight
  for (uint32_t idx = tid; idx < nrow * LPR; idx += BAL_TPB) {
    const uint32_t row = idx / LPR, c = (idx % LPR) * CPL;
    const uint32_t q0 = roff[row], q1 = roff[row + 1];
    ST orr[CPL], oi[CPL];
#pragma unroll
    for (int k = 0; k < CPL; k++) {
      orr[k] = 0;
      oi[k] = 0;
    }
    // an absent target reads row 0 and adds nothing (terms are added in
    // generator order either way)
    auto term = [&](uint32_t q, const VT (&xj)[CPL]) {
      const bool hit = tgt[q] != HASH_EMPTY;
#pragma unroll
      for (int k = 0; k < CPL; k++) {
        if (CPLX) {
          orr[k] += hit ? tcr[q] * (ST)xj[k].x - tci[q] * (ST)xj[k].y : (ST)0;
          oi[k] += hit ? tcr[q] * (ST)xj[k].y + tci[q] * (ST)xj[k].x : (ST)0;
        } else {
          orr[k] += hit ? tcr[q] * (ST)xj[k].x : (ST)0;
          oi[k] += hit ? tcr[q] * (ST)xj[k].y : (ST)0;
        }
      }
    };
    auto gather = [&](uint32_t q, VT (&xj)[CPL]) {
      const uint32_t jj = tgt[q];
      const VT *xr = x + (size_t)(jj != HASH_EMPTY ? jj : 0u) * NV + c;
#pragma unroll
      for (int k = 0; k < CPL; k++) xj[k] = xr[k];
    };
    uint32_t q = q0;
    for (; q + GF <= q1; q += GF) {
      VT xs[GF][CPL];
#pragma unroll
      for (int u = 0; u < GF; u++) gather(q + u, xs[u]);
#pragma unroll
      for (int u = 0; u < GF; u++) term(q + u, xs[u]);
    }
    for (; q + 4 <= q1; q += 4) {
      VT xs[4][CPL];
#pragma unroll
      for (int u = 0; u < 4; u++) gather(q + u, xs[u]);
#pragma unroll
      for (int u = 0; u < 4; u++) term(q + u, xs[u]);
    }
    for (; q < q1; q++) {
      VT xs[CPL];
      gather(q, xs);
      term(q, xs);
    }
#pragma unroll
    for (int k = 0; k < CPL; k++) {
      const size_t i = (size_t)(r0 + row) * NV + c + k;
      const VT xi = x[i];
      const ST yr = rdg[row] * (ST)xi.x + orr[k] * rinv[row];
      const ST yi = rdg[row] * (ST)xi.y + oi[k] * rinv[row];
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
}

template <int KIND, int PH, int LT, typename VT, typename ST, int NV>
__global__ void __launch_bounds__(BAL_TPB) k_matvec_bal_blk(FastArgs A, const VT *__restrict__ x,
                                                             VT *__restrict__ y, int accumulate) {
  bal_blk_body<KIND, PH, LT, VT, ST, NV>(A, bal_chunk(blockIdx.x, A.nchunk, A.xcdg), x, y, accumulate);
}

// Several sectors' block products in one launch (the fused form of a
// multi-sector batch): sector k (arguments M.a[k]) owns the workgroups
// [bstart[k], bstart[k+1]),
// a range padded to a multiple of 8 so that each sector's chunks keep the
// XCD-grouped order of its own launch; padding workgroups exit at once.
// Workgroups of small and large sectors run side by side, and the chip
// drains once instead of once per sector.
// kernel arguments are limited to 4 KB (MultiSec plus the accumulate flag)
static_assert(sizeof(MultiSec) + 16 <= 4096, "fused launch arguments exceed the kernel argument limit");

template <int KIND, int PH, int LT, typename VT, typename ST, int NV>
__global__ void __launch_bounds__(BAL_TPB) k_matvec_bal_blk_multi(MultiSec M, int accumulate) {
  // the sector: a count over all range starts (independent scalar reads of
  // the kernel arguments, not a chain of dependent ones)
  int k = 0;
#pragma unroll
  for (int q = 1; q < MS_MAX; q++) k += (q < M.n && blockIdx.x >= M.bstart[q]) ? 1 : 0;
  const FastArgs &A = M.a[k];
  const uint32_t lb = blockIdx.x - M.bstart[k];
  if (lb >= A.nchunk) return;
  bal_blk_body<KIND, PH, LT, VT, ST, NV>(A, bal_chunk(lb, A.nchunk, A.xcdg), (const VT *)M.x[k], (VT *)M.y[k],
                                          accumulate);
}

template <int KIND, int PH, int LT, int NV>
static void launch_blk_multi(const MultiSec &M, int dtype, int acc, hipStream_t st) {
  const uint32_t nb = M.bstart[M.n];
  if (dtype == EDK_C64)
    k_matvec_bal_blk_multi<KIND, PH, LT, double2, double, NV><<<nb, BAL_TPB, 0, st>>>(M, acc);
  else
    k_matvec_bal_blk_multi<KIND, PH, LT, float2, float, NV><<<nb, BAL_TPB, 0, st>>>(M, acc);
}

template <int KIND, int PH, int NV>
static void launch_blk_multi_lt(const MultiSec &M, int dtype, int acc, hipStream_t st) {
  const FastArgs &A0 = M.a[0];
  if (A0.L == 32 && A0.A == 1) launch_blk_multi<KIND, PH, 32, NV>(M, dtype, acc, st);
  else if (A0.A == 1 && A0.L >= 16 && KIND != EDK_PROJECTED) launch_blk_multi<KIND, PH, 1, NV>(M, dtype, acc, st);
  else launch_blk_multi<KIND, PH, 0, NV>(M, dtype, acc, st);
}

template <int KIND, int PH>
static bool launch_blk_multi_nv(const MultiSec &M, int dtype, int nv, int acc, hipStream_t st) {
  switch (nv) {
    case 2: launch_blk_multi_lt<KIND, PH, 2>(M, dtype, acc, st); return true;
    case 4: launch_blk_multi_lt<KIND, PH, 4>(M, dtype, acc, st); return true;
    case 8: launch_blk_multi_lt<KIND, PH, 8>(M, dtype, acc, st); return true;
    default: return false;
  }
}

// the phase mode (0 real positive, 1 real signs, 2 complex) a sector's block
// launch uses; -1 when it has no block launch
int edk_fast_blk_mode(const FastArgs &A) {
  if (!A.chunk) return -1;
  switch (A.kind) {
    case EDK_TRANSLATION_PARITY: return (A.kp == 1 && A.P == 1) ? 0 : 1;
    case EDK_TRANSLATIONAL: return A.kp == 0 ? 0 : (2 * A.kp == A.len ? 1 : 2);
    case EDK_PROJECTED: return 0;
    default: return -1;
  }
}

bool edk_fast_launch_blk_multi(const MultiSec &M, int dtype, int nv, int acc, hipStream_t st) {
  if (M.n < 1 || M.n > MS_MAX) return false;
  const FastArgs &A0 = M.a[0];
  switch (A0.kind) {
    case EDK_TRANSLATION_PARITY:
      if (edk_fast_blk_mode(A0) == 0) return launch_blk_multi_nv<EDK_TRANSLATION_PARITY, 0>(M, dtype, nv, acc, st);
      return launch_blk_multi_nv<EDK_TRANSLATION_PARITY, 1>(M, dtype, nv, acc, st);
    case EDK_TRANSLATIONAL:
      if (edk_fast_blk_mode(A0) == 0) return launch_blk_multi_nv<EDK_TRANSLATIONAL, 0>(M, dtype, nv, acc, st);
      if (edk_fast_blk_mode(A0) == 1) return launch_blk_multi_nv<EDK_TRANSLATIONAL, 1>(M, dtype, nv, acc, st);
      return launch_blk_multi_nv<EDK_TRANSLATIONAL, 2>(M, dtype, nv, acc, st);
    case EDK_PROJECTED: return launch_blk_multi_nv<EDK_PROJECTED, 0>(M, dtype, nv, acc, st);
    default: return false;
  }
}

template <int KIND, int PH, int LT, int NV>
static void launch_blk(const FastArgs &A, int dtype, void *y, const void *x, int acc, hipStream_t st) {
  if (dtype == EDK_C64)
    k_matvec_bal_blk<KIND, PH, LT, double2, double, NV>
        <<<A.nchunk, BAL_TPB, 0, st>>>(A, (const double2 *)x, (double2 *)y, acc);
  else
    k_matvec_bal_blk<KIND, PH, LT, float2, float, NV>
        <<<A.nchunk, BAL_TPB, 0, st>>>(A, (const float2 *)x, (float2 *)y, acc);
}

template <int KIND, int PH, int NV>
static void launch_blk_lt(const FastArgs &A, int dtype, void *y, const void *x, int acc, hipStream_t st) {
  if (A.L == 32 && A.A == 1) launch_blk<KIND, PH, 32, NV>(A, dtype, y, x, acc, st);
  else if (A.A == 1 && A.L >= 16 && KIND != EDK_PROJECTED) launch_blk<KIND, PH, 1, NV>(A, dtype, y, x, acc, st);
  else launch_blk<KIND, PH, 0, NV>(A, dtype, y, x, acc, st);
}

template <int KIND, int PH>
static bool launch_blk_nv(const FastArgs &A, int dtype, void *y, const void *x, int nv, int acc, hipStream_t st) {
  switch (nv) {
    case 2: launch_blk_lt<KIND, PH, 2>(A, dtype, y, x, acc, st); return true;
    case 4: launch_blk_lt<KIND, PH, 4>(A, dtype, y, x, acc, st); return true;
    case 8: launch_blk_lt<KIND, PH, 8>(A, dtype, y, x, acc, st); return true;
    default: return false;
  }
}

bool edk_fast_launch_blk(const FastArgs &A, int dtype, void *y, const void *x, int nv, int acc, hipStream_t st) {
  if (!A.chunk) return false;
  // phase mode as for one vector: 0 real positive, 1 real signs, 2 complex
  switch (A.kind) {
    case EDK_TRANSLATION_PARITY:
      if (A.kp == 1 && A.P == 1) return launch_blk_nv<EDK_TRANSLATION_PARITY, 0>(A, dtype, y, x, nv, acc, st);
      return launch_blk_nv<EDK_TRANSLATION_PARITY, 1>(A, dtype, y, x, nv, acc, st);
    case EDK_TRANSLATIONAL:
      if (A.kp == 0) return launch_blk_nv<EDK_TRANSLATIONAL, 0>(A, dtype, y, x, nv, acc, st);
      if (2 * A.kp == A.len) return launch_blk_nv<EDK_TRANSLATIONAL, 1>(A, dtype, y, x, nv, acc, st);
      return launch_blk_nv<EDK_TRANSLATIONAL, 2>(A, dtype, y, x, nv, acc, st);
    case EDK_PROJECTED: return launch_blk_nv<EDK_PROJECTED, 0>(A, dtype, y, x, nv, acc, st);
    default: return false;
  }
}

// ---------------------------------------------------------------------------
// Multi-coefficient block kernel: NV operators that share a basis and a
// family structure (the same distances, location masks and flips) but not
// their values -- e.g. XXZ chains of different Delta, or J1-J2 rings of
// different J2, in one sector -- applied to the NV columns of one
// row-interleaved block, column c with the values of operator c.  The
// canonical form and the lookup of every target are computed once for all
// columns; phase 3 gives each row NV lanes, and lane c applies its operator's
// entry value and diagonal and writes its own output vector.  A column's
// arithmetic does not depend on NV or on the other columns, so a vector gets
// the same bits in a block of one as in any mix.
// ---------------------------------------------------------------------------
template <int KIND, int PH, int LT, typename VT, typename ST, int NV>
__global__ void __launch_bounds__(BAL_TPB, NV == 1 ? 6 : 1) k_matvec_bal_mc(FastArgs A, McArgs M, const VT *__restrict__ X) {
  // every product and sum is rounded on its own (no fused multiply-add), so
  // the one-column form below and the NV-column form round a column's terms
  // identically: a vector gets the same bits alone and in any block
#pragma clang fp contract(off)
  constexpr bool CPLX = (PH == 2);
  __shared__ double sR[2 * 64];
  __shared__ double2 sP[PH != 0 ? 64 : 1];
  __shared__ double sV[NV][FAST_MAXFAM * FAST_MAXENT];
  __shared__ double sD[NV][FAST_MAXFAM * 4];
  __shared__ uint32_t tgt[BAL_CAP];
  __shared__ uint8_t ent[BAL_CAP];
  // (position, canonical coefficient) per target; with one column the
  // target's contribution (re, im) instead
  __shared__ double tcr[BAL_CAP], tci[(CPLX || NV == 1) ? BAL_CAP : 1];
  // per-row tables for the NV-lane phase 3; one column keeps them in the row
  // thread's registers (6 instead of 4 workgroups per CU)
  constexpr int NROW = NV == 1 ? 1 : BAL_TPB;
  __shared__ uint32_t roff[NROW], rcnt[NROW];
  __shared__ uint8_t rpop[NROW][FAST_MAXFAM * 4];
  __shared__ ST rinv[NROW];
  __shared__ uint32_t wsum[BAL_TPB / 64];
  constexpr bool RUNS = (LT == 32 || LT == 1) && KIND != EDK_PROJECTED;
  __shared__ uint16_t dq[RUNS ? BAL_CAP : 1];
  __shared__ uint32_t ndef;
  const FastTab &T = *A.tab;
  const int tid = threadIdx.x;
  if (RUNS && tid == 0) ndef = 0;
  if (KIND != EDK_PROJECTED) {
    int nr = (KIND == EDK_TRANSLATION_PARITY ? 2 : 1) * A.len;
    if (tid < nr) sR[tid] = A.Rtab[tid];
    if (PH != 0 && KIND == EDK_TRANSLATIONAL && tid < A.len) sP[tid] = A.ph[tid];
  }
  for (int q = tid; q < NV * FAST_MAXFAM * FAST_MAXENT; q += BAL_TPB)
    sV[q / (FAST_MAXFAM * FAST_MAXENT)][q % (FAST_MAXFAM * FAST_MAXENT)] =
        M.ctab[q / (FAST_MAXFAM * FAST_MAXENT)]->val[q % (FAST_MAXFAM * FAST_MAXENT)];
  if (tid < NV * FAST_MAXFAM * 4) sD[tid / (FAST_MAXFAM * 4)][tid % (FAST_MAXFAM * 4)] =
      M.ctab[tid / (FAST_MAXFAM * 4)]->diag[tid % (FAST_MAXFAM * 4)];
  FastArgs B = A;
  B.Rtab = sR;
  if (PH != 0 && KIND == EDK_TRANSLATIONAL) B.ph = sP;
  const uint32_t cb = bal_chunk(blockIdx.x, A.nchunk, A.xcdg);
  const uint32_t r0 = A.chunk[cb], nrow = A.chunk[cb + 1] - r0;
  // phase 1: row threads count and write their targets; the diagonal's
  // popcounts (per family: 00, 01, 10, 11 pairs) and 1/R_i
  uint32_t s = 0, n = 0;
  const bool own = (uint32_t)tid < nrow;
  if (own) {
    s = A.keys[r0 + tid];
    const uint32_t ns = ~s & A.mask;
    for (int f = 0; f < T.nfam; f++) {
      uint32_t r = rotr_lt<LT>(s, A.L - T.fam[f].d, A.L, A.mask);
      for (int e = 0; e < T.fam[f].nent; e++) n += __popc(ent_mask(T.fam[f].ent[e].locs, s, ns, r, A.mask));
    }
  }
  uint32_t total;
  const uint32_t off = block_scan_u32(n, &total, wsum);
  if (own) {
    const uint32_t ns = ~s & A.mask;
    uint32_t o = off;
    for (int f = 0; f < T.nfam; f++) {
      const int d = T.fam[f].d;
      uint32_t r = rotr_lt<LT>(s, A.L - d, A.L, A.mask);
      uint32_t nr = ~r & A.mask;
      if constexpr (NV > 1) {
        rpop[tid][f * 4 + 0] = (uint8_t)__popc(ns & nr);
        rpop[tid][f * 4 + 1] = (uint8_t)__popc(ns & r);
        rpop[tid][f * 4 + 2] = (uint8_t)__popc(s & nr);
        rpop[tid][f * 4 + 3] = (uint8_t)__popc(s & r);
      }
      for (int e = 0; e < T.fam[f].nent; e++) {
        const int flip = T.fam[f].ent[e].flip;
        uint32_t bits = ent_mask(T.fam[f].ent[e].locs, s, ns, r, A.mask);
        while (bits) {
          int p = __ffs(bits) - 1;
          bits &= bits - 1;
          int p2 = p - d;
          if (LT == 32) p2 &= 31;
          else if (p2 < 0) p2 += A.L;
          uint32_t t = s;
          if (flip & 2) t ^= 1u << p;
          if (flip & 1) t ^= 1u << p2;
          tgt[o] = t;
          ent[o] = (uint8_t)(f * FAST_MAXENT + e);
          o++;
        }
      }
    }
    if constexpr (NV > 1) {
      roff[tid] = off;
      rcnt[tid] = n;
      rinv[tid] = (KIND == EDK_PROJECTED) ? (ST)1 : (ST)(1.0 / A.Rrow[r0 + tid]);
    }
  }
  __syncthreads();
  // phase 2: BAL_K targets per thread -> (position, canonical coefficient)
  uint32_t rep[BAL_K], jj[BAL_K];
  bool ok[BAL_K];
  double cfr[BAL_K], cfi[CPLX ? BAL_K : 1];
  if constexpr (RUNS) {
#pragma unroll 1
    for (int q = 0; q < BAL_K; q++) {
      const uint32_t j = tid + BAL_TPB * q;
      if (j < total) {
        uint32_t r;
        double cr = 0.0, ci = 0.0;
        const int st = canon_runs<KIND, PH, LT>(B, tgt[j], r, cr, ci);
        if (st > 0) {
          tgt[j] = r;
          tcr[j] = cr;
          if (CPLX) tci[j] = ci;
        } else if (st == 0) {
          ent[j] = 0xff;
        } else {
          dq[atomicAdd(&ndef, 1u)] = (uint16_t)j;
        }
      }
    }
    __syncthreads();
    const uint32_t nd = ndef;
    for (uint32_t d = tid; d < nd; d += BAL_TPB) {
      const uint32_t j = dq[d];
      uint32_t r = 0;
      double cr = 0.0, ci = 0.0;
      const bool okb = (LT == 32) ? fast_canon32<KIND, PH>(B, tgt[j], r, cr, ci)
                                  : fast_canon_w<KIND, PH>(B, tgt[j], r, cr, ci);
      if (okb) {
        tgt[j] = r;
        tcr[j] = cr;
        if (CPLX) tci[j] = ci;
      } else {
        ent[j] = 0xff;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BAL_K; q++) {
      const uint32_t j = tid + BAL_TPB * q;
      ok[q] = j < total && ent[j] != 0xff;
      rep[q] = ok[q] ? tgt[j] : 0u;
      cfr[q] = ok[q] ? tcr[j] : 0.0;
      if (CPLX) cfi[q] = ok[q] ? tci[j] : 0.0;
    }
  } else {
#pragma unroll
    for (int q = 0; q < BAL_K; q++) {
      const uint32_t j = tid + BAL_TPB * q;
      ok[q] = j < total;
      uint32_t t = ok[q] ? tgt[j] : 0u;
      double cr = 0.0, ci = 0.0;
      rep[q] = 0;
      if (ok[q]) {
        if (LT == 32) ok[q] = fast_canon32<KIND, PH>(B, t, rep[q], cr, ci);
        else ok[q] = fast_canon<KIND, PH>(B, t, rep[q], cr, ci);
      }
      cfr[q] = cr;
      if (CPLX) cfi[q] = ci;
    }
  }
  region_lookup_many<BAL_K>(A, rep, ok, jj);
  if constexpr (NV == 1) {
    // one column: the gathers of a lane's targets issued together (as in
    // k_matvec_bal), contributions to LDS; an absent target reads x[0] and
    // is dropped
    VT xg[BAL_K];
#pragma unroll
    for (int q = 0; q < BAL_K; q++) xg[q] = X[jj[q] != HASH_EMPTY ? jj[q] : 0u];
#pragma unroll
    for (int q = 0; q < BAL_K; q++) {
      const uint32_t j = tid + BAL_TPB * q;
      if (j < total) {
        tgt[j] = jj[q];
        const double v = jj[q] != HASH_EMPTY ? sV[0][ent[j]] : 0.0;
        const ST cr = (ST)(v * cfr[q]);
        ST pr, pi;
        if (CPLX) {
          const ST ci = (ST)(v * cfi[q]);
          pr = cr * (ST)xg[q].x - ci * (ST)xg[q].y;
          pi = cr * (ST)xg[q].y + ci * (ST)xg[q].x;
        } else {
          pr = cr * (ST)xg[q].x;
          pi = cr * (ST)xg[q].y;
        }
        tcr[j] = (double)pr;
        tci[j] = (double)pi;
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < BAL_K; q++) {
      const uint32_t j = tid + BAL_TPB * q;
      if (j < total) {
        tgt[j] = jj[q];
        tcr[j] = cfr[q];
        if (CPLX) tci[j] = cfi[q];
      }
    }
  }
  __syncthreads();
  if constexpr (NV == 1) {
    // phase 3, one column: row threads sum their contributions in order
    if (own && M.y[0]) {
      VT *yc = (VT *)M.y[0];
      const uint32_t q0 = off, q1 = off + n;
      // the row's popcounts and 1/R_i again (the values the NV-lane form
      // keeps in LDS)
      const uint32_t ns = ~s & A.mask;
      ST dg = 0;
      for (int f = 0; f < T.nfam; f++) {
        const uint32_t r = rotr_lt<LT>(s, A.L - T.fam[f].d, A.L, A.mask), nr = ~r & A.mask;
        const uint8_t p0 = (uint8_t)__popc(ns & nr), p1 = (uint8_t)__popc(ns & r), p2 = (uint8_t)__popc(s & nr),
                      p3 = (uint8_t)__popc(s & r);
        dg += (ST)(sD[0][f * 4 + 0] * p0 + sD[0][f * 4 + 1] * p1 + sD[0][f * 4 + 2] * p2 + sD[0][f * 4 + 3] * p3);
      }
      const ST rinv0 = (KIND == EDK_PROJECTED) ? (ST)1 : (ST)(1.0 / A.Rrow[r0 + tid]);
      ST orr = 0, oi = 0;
      for (uint32_t q = q0; q < q1; q++) {
        if (tgt[q] == HASH_EMPTY) continue;
        orr += (ST)tcr[q];
        oi += (ST)tci[q];
      }
      const VT xi = X[r0 + tid];
      VT out;
      out.x = dg * (ST)xi.x + orr * rinv0;
      out.y = dg * (ST)xi.y + oi * rinv0;
      yc[r0 + tid] = out;
    }
    return;
  }
  // phase 3: NV lanes per row; lane c sums its column in generator order with
  // its operator's values, adds its diagonal and writes its output
  for (uint32_t idx = tid; idx < nrow * NV; idx += BAL_TPB) {
    const uint32_t row = idx / NV, c = idx % NV;
    VT *yc = (VT *)M.y[c];
    if (!yc) continue;
    const uint32_t q0 = roff[row], q1 = q0 + rcnt[row];
    ST dg = 0;
    for (int f = 0; f < T.nfam; f++)
      dg += (ST)(sD[c][f * 4 + 0] * rpop[row][f * 4 + 0] + sD[c][f * 4 + 1] * rpop[row][f * 4 + 1] +
                 sD[c][f * 4 + 2] * rpop[row][f * 4 + 2] + sD[c][f * 4 + 3] * rpop[row][f * 4 + 3]);
    ST orr = 0, oi = 0;
    // four gathers in flight per lane; an absent target reads row 0 and adds
    // nothing.  Each product is formed as in the one-column path before it is
    // added, in generator order, so a column's result does not depend on NV.
    auto gather = [&](uint32_t q) {
      const uint32_t jq = tgt[q];
      return X[(size_t)(jq != HASH_EMPTY ? jq : 0u) * NV + c];
    };
    auto term = [&](uint32_t q, const VT &xj) {
      const bool hit = tgt[q] != HASH_EMPTY;
      const double v = sV[c][hit ? ent[q] : 0];
      const ST cr = (ST)(v * tcr[q]);
      ST pr, pi;
      if (CPLX) {
        const ST ci = (ST)(v * tci[q]);
        pr = cr * (ST)xj.x - ci * (ST)xj.y;
        pi = cr * (ST)xj.y + ci * (ST)xj.x;
      } else {
        pr = cr * (ST)xj.x;
        pi = cr * (ST)xj.y;
      }
      orr += hit ? pr : (ST)0;
      oi += hit ? pi : (ST)0;
    };
    uint32_t q = q0;
    for (; q + 4 <= q1; q += 4) {
      const VT x0 = gather(q), x1 = gather(q + 1), x2 = gather(q + 2), x3 = gather(q + 3);
      term(q, x0);
      term(q + 1, x1);
      term(q + 2, x2);
      term(q + 3, x3);
    }
    for (; q < q1; q++) term(q, gather(q));
    const VT xi = X[(size_t)(r0 + row) * NV + c];
    VT out;
    out.x = dg * (ST)xi.x + orr * rinv[row];
    out.y = dg * (ST)xi.y + oi * rinv[row];
    yc[r0 + row] = out;
  }
}

template <int KIND, int PH, int LT, int NV>
static void launch_mc(const FastArgs &A, const McArgs &M, int dtype, const void *X, hipStream_t st) {
  if (dtype == EDK_C64)
    k_matvec_bal_mc<KIND, PH, LT, double2, double, NV><<<A.nchunk, BAL_TPB, 0, st>>>(A, M, (const double2 *)X);
  else
    k_matvec_bal_mc<KIND, PH, LT, float2, float, NV><<<A.nchunk, BAL_TPB, 0, st>>>(A, M, (const float2 *)X);
}

template <int KIND, int PH, int NV>
static void launch_mc_lt(const FastArgs &A, const McArgs &M, int dtype, const void *X, hipStream_t st) {
  if (A.L == 32 && A.A == 1) launch_mc<KIND, PH, 32, NV>(A, M, dtype, X, st);
  else if (A.A == 1 && A.L >= 16 && KIND != EDK_PROJECTED) launch_mc<KIND, PH, 1, NV>(A, M, dtype, X, st);
  else launch_mc<KIND, PH, 0, NV>(A, M, dtype, X, st);
}

template <int KIND, int PH>
static bool launch_mc_nv(const FastArgs &A, const McArgs &M, int dtype, const void *X, int nv, hipStream_t st) {
  switch (nv) {
    case 1: launch_mc_lt<KIND, PH, 1>(A, M, dtype, X, st); return true;
    case 2: launch_mc_lt<KIND, PH, 2>(A, M, dtype, X, st); return true;
    case 4: launch_mc_lt<KIND, PH, 4>(A, M, dtype, X, st); return true;
    case 8: launch_mc_lt<KIND, PH, 8>(A, M, dtype, X, st); return true;
    default: return false;
  }
}

bool edk_fast_launch_mc(const FastArgs &A, const McArgs &M, int dtype, const void *X, int nv, hipStream_t st) {
  if (!A.chunk) return false;
  switch (A.kind) {
    case EDK_TRANSLATION_PARITY:
      if (A.kp == 1 && A.P == 1) return launch_mc_nv<EDK_TRANSLATION_PARITY, 0>(A, M, dtype, X, nv, st);
      return launch_mc_nv<EDK_TRANSLATION_PARITY, 1>(A, M, dtype, X, nv, st);
    case EDK_TRANSLATIONAL:
      if (A.kp == 0) return launch_mc_nv<EDK_TRANSLATIONAL, 0>(A, M, dtype, X, nv, st);
      if (2 * A.kp == A.len) return launch_mc_nv<EDK_TRANSLATIONAL, 1>(A, M, dtype, X, nv, st);
      return launch_mc_nv<EDK_TRANSLATIONAL, 2>(A, M, dtype, X, nv, st);
    case EDK_PROJECTED: return launch_mc_nv<EDK_PROJECTED, 0>(A, M, dtype, X, nv, st);
    default: return false;
  }
}

// X[i * nv + c] = xs[c][i]: nv plain vectors into one row-interleaved block
template <typename VT>
__global__ void __launch_bounds__(256) k_gather_cols(ColPtrs G, int nv, int64_t D, VT *__restrict__ X) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= D * nv) return;
  const int64_t i = t / nv;
  const int c = (int)(t - i * nv);
  X[t] = ((const VT *)G.p[c])[i];
}

int edk_gather_cols(const ColPtrs &G, int nv, int64_t D, int dtype, void *X, hipStream_t st) {
  for (int c = 0; c < nv; c++)
    if (!G.p[c]) return edk_fail(EDK_ERR_ARG, "edk_gather_cols: null column");
  const unsigned g = (unsigned)((D * nv + 255) / 256);
  if (dtype == EDK_C64) k_gather_cols<double2><<<g, 256, 0, st>>>(G, nv, D, (double2 *)X);
  else k_gather_cols<float2><<<g, 256, 0, st>>>(G, nv, D, (float2 *)X);
  EDK_HIP(hipGetLastError());
  return EDK_OK;
}

