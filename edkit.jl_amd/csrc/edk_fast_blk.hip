// edk_fast_blk.hip -- specialised gather kernel for blocks of NV vectors
// (A10: mul!(Y, H, X), Operator.jl:322-328, 382-394).
#include "edk_fast.h"

// Block of NV vectors, row-interleaved (element (i, c) at x[i*NV + c]): the
// same chunks and phases as k_matvec_bal, but phase 2 keeps each target's
// (position, coefficient) in LDS, and phase 3 gives every row NV consecutive
// threads, one per vector, so each representative lookup serves all NV
// columns and the NV gathered values of a target form one contiguous segment.
template <int KIND, int PH, int LT, typename VT, typename ST, int NV>
__global__ void __launch_bounds__(BAL_TPB) k_matvec_bal_blk(FastArgs A, const VT *__restrict__ x,
                                                             VT *__restrict__ y, int accumulate) {
  constexpr bool CPLX = (PH == 2);
  __shared__ double sR[2 * 64];
  __shared__ double2 sP[PH != 0 ? 64 : 1];
  __shared__ double sV[FAST_MAXFAM * FAST_MAXENT];
  __shared__ uint32_t tgt[BAL_CAP];
  __shared__ uint8_t ent[BAL_CAP];
  __shared__ ST tcr[BAL_CAP], tci[CPLX ? BAL_CAP : 1];
  __shared__ uint32_t roff[BAL_TPB], rcnt[BAL_TPB];
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
  const uint32_t r0 = A.chunk[blockIdx.x], nrow = A.chunk[blockIdx.x + 1] - r0;
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
    rcnt[tid] = n;
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
  // phase 3: NV threads per row, targets summed in generator order per column
  for (uint32_t idx = tid; idx < nrow * NV; idx += BAL_TPB) {
    const uint32_t row = idx / NV, c = idx % NV;
    const uint32_t q0 = roff[row], q1 = q0 + rcnt[row];
    ST orr = 0, oi = 0;
    for (uint32_t q = q0; q < q1; q++) {
      const uint32_t jj = tgt[q];
      if (jj == HASH_EMPTY) continue;
      const VT xj = x[(size_t)jj * NV + c];
      if (CPLX) {
        orr += tcr[q] * (ST)xj.x - tci[q] * (ST)xj.y;
        oi += tcr[q] * (ST)xj.y + tci[q] * (ST)xj.x;
      } else {
        orr += tcr[q] * (ST)xj.x;
        oi += tcr[q] * (ST)xj.y;
      }
    }
    const size_t i = (size_t)(r0 + row) * NV + c;
    const VT xi = x[i];
    const ST yr = rdg[row] * (ST)xi.x + orr * rinv[row];
    const ST yi = rdg[row] * (ST)xi.y + oi * rinv[row];
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
