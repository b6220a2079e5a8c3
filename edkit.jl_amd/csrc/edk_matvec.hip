// edk_matvec.hip -- Operator upload and the atomics-free gather matvec.
//
// The reference applies H column by column and scatters (colmn!,
// Operator.jl:426-465): for column j, every nonzero (row, val) of every local
// term at loc_t(a_j) rewrites the term's digits and adds
//     y[pos] += (x_j / R_j) * (phase * R_pos) * val.
// The device computes each output row instead (gather): for row i and term t
// it walks row loc_t(a_i) of M_t, i.e. the entries M_t[loc, c], builds
// b' = a_i with the term's sites set to c, canonicalises b' to (rep, phase),
// finds rep's position j in the sorted table and adds
//     y_i += M_t[loc, c] * conj(phase) * (R_j / R_i) * x_j ,
// which is the (i, j) element of the same sector matrix when the operator
// commutes with the symmetry group.  Diagonal local entries (c == loc) give
// b' = a_i and contribute M_t[loc, loc] * x_i without any lookup.  Each row is
// written once by one lane: no atomics, and the summation order per row is
// fixed, so results are bitwise reproducible run to run.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "edk_matvec.h"

// ---------------------------------------------------------------------------
// generic kernel: one lane per row, any basis kind, any local terms.  NV
// columns [col, col + NV) of a row-interleaved block share each target's
// canonical form and lookup.
// ---------------------------------------------------------------------------
// STAGED: the local-term CSR tables (every distinct matrix) are copied into
// LDS first; the host picks it when they fit GEN_LDS_* entries.
#define GEN_LDS_MO 64
#define GEN_LDS_RP 512
#define GEN_LDS_NZ 512
template <typename VT, typename ST, int NV, bool STAGED>
__global__ void __launch_bounds__(256) k_matvec_generic(MatvecArgs A, const VT *__restrict__ x, VT *__restrict__ y,
                                                        int64_t nb, int64_t col, int accumulate) {
  __shared__ int s_mo[STAGED ? GEN_LDS_MO : 1], s_rp[STAGED ? GEN_LDS_RP : 1], s_col[STAGED ? GEN_LDS_NZ : 1];
  __shared__ double2 s_val[STAGED ? GEN_LDS_NZ : 1];
  TermsDev T = A.T;
  if constexpr (STAGED) {
    for (int k = threadIdx.x; k < T.nmo; k += blockDim.x) s_mo[k] = T.mrow_off[k];
    for (int k = threadIdx.x; k < T.nrp; k += blockDim.x) s_rp[k] = T.rowptr[k];
    for (int k = threadIdx.x; k < T.nnz; k += blockDim.x) {
      s_col[k] = T.col[k];
      s_val[k] = T.val[k];
    }
    __syncthreads();
    T.mrow_off = s_mo;
    T.rowptr = s_rp;
    T.col = s_col;
    T.val = s_val;
  }
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.D) return;
  uint64_t fm = (1ull << A.w) - 1;
  uint64_t s = A.kind == EDK_TENSOR ? edk_content_to_packed_d((uint64_t)i, A.L, A.base, A.w) : A.S.keys[i];
  ST dr = 0, di = 0, orr[NV], oi[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) orr[c] = oi[c] = 0;
  for (int t = 0; t < T.nterms; t++) {
    const int *td = T.tdesc + t * EDK_TDESC;
    int ns = td[0], m = td[1];
    int loc = 0;
    for (int q = 0; q < ns; q++) loc = loc * A.base + (int)((s >> td[2 + q]) & fm);
    const int *rp = T.rowptr + T.mrow_off[m];
    int e0 = rp[loc], e1 = rp[loc + 1];
    for (int e = e0; e < e1; e++) {
      int c = T.col[e];
      double2 v = T.val[e];
      if (c == loc) {
        dr += (ST)v.x;
        di += (ST)v.y;
        continue;
      }
      uint64_t s2 = s;
      int64_t delta = 0;
      int cc = c;
      for (int q = ns - 1; q >= 0; q--) {
        // the local column's digits, last site first (base 2: its bits)
        uint64_t nd;
        if (A.base == 2) {
          nd = (uint64_t)(cc & 1);
          cc >>= 1;
        } else {
          nd = (uint64_t)(cc % A.base);
          cc /= A.base;
        }
        int sh = td[2 + q];
        uint64_t od = (s2 >> sh) & fm;
        s2 = (s2 & ~(fm << sh)) | (nd << sh);
        delta += ((int64_t)nd - (int64_t)od) * (int64_t)T.tpw[t * EDK_MAXSITES + q];
      }
      int64_t j;
      double cr = 1.0, ci = 0.0;
      if (A.kind == EDK_TENSOR) {
        j = i + delta;
      } else if (A.kind == EDK_PROJECTED) {
        j = edk_search(A.S, s2);
        if (j < 0) continue;
      } else {
        CanonOut o = edk_canon(A.C, s2);
        if (A.orbit_norms && !o.ok) continue;
        j = edk_search(A.S, o.rep);
        if (j < 0) continue;
        double Rj = A.orbit_norms ? o.R : A.R[j];
        cr = o.phase.x * Rj;   // conj(phase) * R_j
        ci = -o.phase.y * Rj;
      }
      // (v * coef) * xj
      ST ar = (ST)(v.x * cr - v.y * ci), ai = (ST)(v.x * ci + v.y * cr);
      const VT *xj = x + j * nb + col;
#pragma unroll
      for (int c = 0; c < NV; c++) {
        VT xv = xj[c];
        orr[c] += ar * (ST)xv.x - ai * (ST)xv.y;
        oi[c] += ar * (ST)xv.y + ai * (ST)xv.x;
      }
    }
  }
  ST invR = (A.kind == EDK_TENSOR || A.kind == EDK_PROJECTED) ? (ST)1 : (ST)(1.0 / A.R[i]);
#pragma unroll
  for (int c = 0; c < NV; c++) {
    VT xi = x[i * nb + col + c];
    ST yr = dr * (ST)xi.x - di * (ST)xi.y + orr[c] * invR;
    ST yi = dr * (ST)xi.y + di * (ST)xi.x + oi[c] * invR;
    VT out;
    if (accumulate) {
      VT y0 = y[i * nb + col + c];
      out.x = y0.x + yr;
      out.y = y0.y + yi;
    } else {
      out.x = yr;
      out.y = yi;
    }
    y[i * nb + col + c] = out;
  }
}

template <typename VT, typename ST, bool STAGED>
static void launch_generic_t(const MatvecArgs &a, const VT *x, VT *y, int64_t nb, int accumulate, hipStream_t st) {
  unsigned grid = (unsigned)((a.D + 255) / 256);
  for (int64_t c = 0; c < nb;) {
    const int64_t left = nb - c;
    if (left >= 8) {
      k_matvec_generic<VT, ST, 8, STAGED><<<grid, 256, 0, st>>>(a, x, y, nb, c, accumulate);
      c += 8;
    } else if (left >= 4) {
      k_matvec_generic<VT, ST, 4, STAGED><<<grid, 256, 0, st>>>(a, x, y, nb, c, accumulate);
      c += 4;
    } else if (left >= 2) {
      k_matvec_generic<VT, ST, 2, STAGED><<<grid, 256, 0, st>>>(a, x, y, nb, c, accumulate);
      c += 2;
    } else {
      k_matvec_generic<VT, ST, 1, STAGED><<<grid, 256, 0, st>>>(a, x, y, nb, c, accumulate);
      c += 1;
    }
  }
}

template <typename VT, typename ST>
static void launch_generic(const MatvecArgs &a, const VT *x, VT *y, int64_t nb, int accumulate, hipStream_t st) {
  const TermsDev &T = a.T;
  if (T.nmo <= GEN_LDS_MO && T.nrp <= GEN_LDS_RP && T.nnz <= GEN_LDS_NZ)
    launch_generic_t<VT, ST, true>(a, x, y, nb, accumulate, st);
  else
    launch_generic_t<VT, ST, false>(a, x, y, nb, accumulate, st);
}

// ---------------------------------------------------------------------------
// generic tile kernel (one vector).  In the one-lane-per-row kernel a lane
// whose row has no target at a term idles while the others canonicalise, so
// half the lanes idle for a two-site spin exchange.  Here the targets are
// compacted into LDS first and every lane of the block canonicalises them.
// One block per chunk of whole rows holding at most GT_CAP off-diagonal local
// entries (chunk: first row of each chunk, built once per operator):
//   phase 1: GT_Q lanes per row walk a share of its terms each; diagonal
//            entries stay in registers, off-diagonal images b' and their
//            entry ids go to LDS at the lane's offset from a block scan, so
//            each row's targets lie in term and entry order;
//   phase 2: lane tid takes targets tid + GT_TPB q: their canonical forms one
//            at a time, then their GT_K lookups and x gathers together;
//   phase 3: each row's first lane adds its shares' diagonal sums, sums the
//            row's slice in order, scales by 1/R_i and writes y.
// A row's off-diagonal sum runs in the row kernel's order; no atomics.
// ---------------------------------------------------------------------------
#define GT_TPB 256
#define GT_K 4
#define GT_CAP (GT_TPB * GT_K)
#define GT_NONE (~0ull)
#define GT_ABSENT 0xffff  // entry-id marker: representative not in the table
#define GT_Q 4  // lanes per row in phase 1

__device__ __forceinline__ uint32_t gt_block_scan(uint32_t v, uint32_t *total, uint32_t *wsum) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t xx = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t yv = __shfl_up(xx, o, 64);
    if (lane >= o) xx += yv;
  }
  if (lane == 63) wsum[wid] = xx;
  __syncthreads();
  uint32_t base = 0, tot = 0;
  for (int q = 0; q < GT_TPB / 64; q++) {
    if (q < wid) base += wsum[q];
    tot += wsum[q];
  }
  *total = tot;
  return base + xx - v;
}

// NV = 0: one vector (x, y plain).  NV >= 1: columns [col, col + NV) of a
// row-interleaved block of nb columns; phase 2 keeps each target's position and
// coefficient in LDS and phase 3 gives each row NV lanes, one per column, that
// walk the row's targets in order (each lookup serves all NV columns).
template <typename VT, typename ST, int NV>
__global__ void __launch_bounds__(GT_TPB, 6) k_matvec_gtile(MatvecArgs A, const uint32_t *__restrict__ chunk,
                                                         const VT *__restrict__ x, VT *__restrict__ y,
                                                         int accumulate, int64_t nb, int64_t col) {
  __shared__ uint32_t s_off[NV ? GT_TPB / GT_Q : 1], s_n[NV ? GT_TPB / GT_Q : 1];
  __shared__ VT s_dg[NV ? GT_TPB / GT_Q : 1];
  __shared__ uint64_t word[GT_CAP];  // image b', then its representative (blocks: its position)
  __shared__ uint16_t ent[GT_CAP];   // entry id into T.col / T.val, or GT_ABSENT
  __shared__ VT contr[GT_CAP];       // coefficient, then contribution
  __shared__ uint32_t wsum[GT_TPB / 64];
  const TermsDev &T = A.T;
  const int tid = threadIdx.x;
  // GT_Q lanes per row (consecutive lanes of one wave), each walking one
  // share of the terms; the block scan in lane order then lays each row's
  // targets out in term order
  const uint32_t r0 = chunk[blockIdx.x], nrow = chunk[blockIdx.x + 1] - r0;
  const uint32_t row = (uint32_t)tid / GT_Q;
  const int qtr = tid % GT_Q;
  const bool own = row < nrow;
  const int t0 = (qtr * T.nterms) / GT_Q, t1 = ((qtr + 1) * T.nterms) / GT_Q;
  const uint64_t fm = (1ull << A.w) - 1;
  uint64_t s = 0;
  uint32_t n = 0;
  ST dr = 0, di = 0;
  if (own) {
    s = A.S.keys[(int64_t)r0 + row];
    for (int t = t0; t < t1; t++) {
      const int *td = T.tdesc + t * EDK_TDESC;
      const int ns = td[0], m = td[1];
      int loc = 0;
      for (int q = 0; q < ns; q++) loc = loc * A.base + (int)((s >> td[2 + q]) & fm);
      const int *rp = T.rowptr + T.mrow_off[m];
      for (int e = rp[loc]; e < rp[loc + 1]; e++) {
        if (T.col[e] == loc) {
          const double2 v = T.val[e];
          dr += (ST)v.x;
          di += (ST)v.y;
        } else {
          n++;
        }
      }
    }
  }
  uint32_t total;
  const uint32_t off = gt_block_scan(n, &total, wsum);
  if (own) {
    uint32_t o = off;
    for (int t = t0; t < t1; t++) {
      const int *td = T.tdesc + t * EDK_TDESC;
      const int ns = td[0], m = td[1];
      int loc = 0;
      for (int q = 0; q < ns; q++) loc = loc * A.base + (int)((s >> td[2 + q]) & fm);
      const int *rp = T.rowptr + T.mrow_off[m];
      for (int e = rp[loc]; e < rp[loc + 1]; e++) {
        int cc = T.col[e];
        if (cc == loc) continue;
        uint64_t s2 = s;
        for (int q = ns - 1; q >= 0; q--) {
          uint64_t nd;
          if (A.base == 2) {
            nd = (uint64_t)(cc & 1);
            cc >>= 1;
          } else {
            nd = (uint64_t)(cc % A.base);
            cc /= A.base;
          }
          const int sh = td[2 + q];
          s2 = (s2 & ~(fm << sh)) | (nd << sh);
        }
        word[o] = s2;
        ent[o] = (uint16_t)e;
        o++;
      }
    }
  }
  __syncthreads();
  // phase 2a: canonical forms, one target at a time: the coefficient into
  // the target's contribution slot, the representative over its image, and
  // GT_ABSENT over the entry id of an image whose representative cannot be
  // in the table.  Each slot is read and rewritten by its own lane only.
#pragma unroll 1
  for (int q = 0; q < GT_K; q++) {
    const uint32_t j = tid + GT_TPB * q;
    if (j < total) {
      const uint64_t t = word[j];
      const double2 v = T.val[ent[j]];
      double cr = 1.0, ci = 0.0;
      uint64_t rep = t;
      bool ok = true;
      if (A.kind != EDK_PROJECTED) {
        const CanonOut o = edk_canon(A.C, t);
        rep = o.rep;
        ok = !(A.orbit_norms && !o.ok);
        const double Rj = A.orbit_norms ? o.R : 1.0;  // uploaded tables: R_j after the lookup
        cr = o.phase.x * Rj;                           // conj(phase) * R_j
        ci = -o.phase.y * Rj;
      }
      VT c;
      c.x = (ST)(v.x * cr - v.y * ci);
      c.y = (ST)(v.x * ci + v.y * cr);
      contr[j] = c;
      word[j] = rep;
      if (!ok) ent[j] = GT_ABSENT;
    }
  }
  // phase 2b: the lane's GT_K lookups together: directory slots, then
  // S.nprobe halving steps (enough for the fullest slot) with the GT_K probes
  // of a step in flight at once, then the equality check.  Inactive lookups
  // read index 0 and are dropped by selection.
  const SearchDev &S = A.S;
  uint64_t key[GT_K];
  uint32_t lo[GT_K], nn[GT_K];
  bool act[GT_K];
#pragma unroll
  for (int q = 0; q < GT_K; q++) {
    const uint32_t j = tid + GT_TPB * q;
    act[q] = j < total && ent[j] != GT_ABSENT;
    key[q] = act[q] ? word[j] : 0;
    const uint64_t b = key[q] >> S.dshift;
    act[q] = act[q] && b < (uint64_t)S.ndir;
    const uint64_t bb = act[q] ? b : 0;
    lo[q] = S.dir[bb];
    nn[q] = act[q] ? S.dir[bb + 1] - lo[q] : 0;
  }
  for (int it = 0; it < S.nprobe; it++) {
    uint64_t kv[GT_K];
#pragma unroll
    for (int q = 0; q < GT_K; q++) kv[q] = S.keys[nn[q] ? lo[q] + (nn[q] >> 1) : 0];
#pragma unroll
    for (int q = 0; q < GT_K; q++) {
      const uint32_t half = nn[q] >> 1;
      if (nn[q]) {
        if (kv[q] < key[q]) {
          lo[q] += half + 1;
          nn[q] -= half + 1;
        } else {
          nn[q] = half;
        }
      }
    }
  }
  uint64_t p[GT_K];
#pragma unroll
  for (int q = 0; q < GT_K; q++) {
    const bool in = act[q] && (int64_t)lo[q] < S.D;
    const uint64_t kv = S.keys[in ? lo[q] : 0];
    p[q] = (in && kv == key[q]) ? (uint64_t)lo[q] : GT_NONE;
  }
  // the row's shares: counts and diagonal sums (share order), at its first lane
  const int lead = (tid & 63) - qtr;  // the row's first lane within the wave
  uint32_t nr = 0;
  ST dsr = 0, dsi = 0;
#pragma unroll
  for (int q = 0; q < GT_Q; q++) {
    nr += __shfl(n, lead + q, 64);
    dsr += __shfl(dr, lead + q, 64);
    dsi += __shfl(di, lead + q, 64);
  }
  if constexpr (NV == 0) {
    // phase 2c: the gathers in flight together (and, for uploaded tables,
    // R_j); an absent target reads x[0] and is dropped by selection
    VT xg[GT_K];
    ST rj[GT_K];
#pragma unroll
    for (int q = 0; q < GT_K; q++) {
      const uint64_t pq = p[q] != GT_NONE ? p[q] : 0;
      xg[q] = x[pq];
      rj[q] = (A.orbit_norms || A.kind == EDK_PROJECTED) ? (ST)1 : (ST)A.R[pq];
    }
#pragma unroll
    for (int q = 0; q < GT_K; q++) {
      const uint32_t j = tid + GT_TPB * q;
      if (j < total) {
        VT c = contr[j];
        if (!A.orbit_norms && A.kind != EDK_PROJECTED) {
          c.x *= rj[q];
          c.y *= rj[q];
        }
        VT cv;
        cv.x = c.x * (ST)xg[q].x - c.y * (ST)xg[q].y;
        cv.y = c.x * (ST)xg[q].y + c.y * (ST)xg[q].x;
        if (p[q] == GT_NONE) cv.x = cv.y = 0;
        contr[j] = cv;
      }
    }
    __syncthreads();
    // phase 3: each row's first lane sums its slice in order
    if (own && qtr == 0) {
      const int64_t i = (int64_t)r0 + row;
      const VT xi = x[i];
      const ST invR = A.kind != EDK_PROJECTED ? (ST)(1.0 / A.R[i]) : (ST)1;
      ST orr = 0, oi = 0;
      for (uint32_t q = off; q < off + nr; q++) {
        const VT cv = contr[q];
        orr += cv.x;
        oi += cv.y;
      }
      const ST yr = dsr * (ST)xi.x - dsi * (ST)xi.y + orr * invR;
      const ST yi = dsr * (ST)xi.y + dsi * (ST)xi.x + oi * invR;
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
  } else {
    // phase 2c: positions over the representatives (own slots); uploaded
    // tables scale the coefficient by R_j now
#pragma unroll
    for (int q = 0; q < GT_K; q++) {
      const uint32_t j = tid + GT_TPB * q;
      if (j < total) {
        word[j] = p[q];
        if (!A.orbit_norms && A.kind != EDK_PROJECTED && p[q] != GT_NONE) {
          const ST r = (ST)A.R[p[q]];
          VT c = contr[j];
          c.x *= r;
          c.y *= r;
          contr[j] = c;
        }
      }
    }
    if (own && qtr == 0) {
      s_off[row] = off;
      s_n[row] = nr;
      VT d;
      d.x = dsr;
      d.y = dsi;
      s_dg[row] = d;
    }
    __syncthreads();
    // phase 3: NV lanes per row, one per column, each walking the row's
    // targets in order; the NV values of a target are contiguous in x
    for (uint32_t idx = tid; idx < nrow * NV; idx += GT_TPB) {
      const uint32_t rw = idx / NV, c = idx % NV;
      const int64_t i = (int64_t)r0 + rw;
      const uint32_t o = s_off[rw], cnt = s_n[rw];
      ST orr = 0, oi = 0;
      // four gathers in flight per lane; an absent target reads row 0 and
      // adds nothing (terms in order either way)
      auto gather = [&](uint32_t q) {
        const uint64_t pq = word[q];
        return x[(int64_t)(pq != GT_NONE ? pq : 0) * nb + col + c];
      };
      auto term = [&](uint32_t q, const VT &xv) {
        const bool hit = word[q] != GT_NONE;
        const VT cf = contr[q];
        orr += hit ? cf.x * (ST)xv.x - cf.y * (ST)xv.y : (ST)0;
        oi += hit ? cf.x * (ST)xv.y + cf.y * (ST)xv.x : (ST)0;
      };
      uint32_t q = o;
      for (; q + 4 <= o + cnt; q += 4) {
        const VT x0 = gather(q), x1 = gather(q + 1), x2 = gather(q + 2), x3 = gather(q + 3);
        term(q, x0);
        term(q + 1, x1);
        term(q + 2, x2);
        term(q + 3, x3);
      }
      for (; q < o + cnt; q++) term(q, gather(q));
      const VT dg = s_dg[rw];
      const VT xi = x[i * nb + col + c];
      const ST invR = A.kind != EDK_PROJECTED ? (ST)(1.0 / A.R[i]) : (ST)1;
      const ST yr = dg.x * (ST)xi.x - dg.y * (ST)xi.y + orr * invR;
      const ST yi = dg.x * (ST)xi.y + dg.y * (ST)xi.x + oi * invR;
      VT out;
      if (accumulate) {
        const VT y0 = y[i * nb + col + c];
        out.x = y0.x + yr;
        out.y = y0.y + yi;
      } else {
        out.x = yr;
        out.y = yi;
      }
      y[i * nb + col + c] = out;
    }
  }
}

// off-diagonal local entries per row (targets before the lookup), for chunking
__global__ void __launch_bounds__(256) k_row_entries(MatvecArgs A, uint32_t *cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.D) return;
  const TermsDev &T = A.T;
  const uint64_t fm = (1ull << A.w) - 1;
  const uint64_t s = A.S.keys[i];
  uint32_t n = 0;
  for (int t = 0; t < T.nterms; t++) {
    const int *td = T.tdesc + t * EDK_TDESC;
    const int ns = td[0], m = td[1];
    int loc = 0;
    for (int q = 0; q < ns; q++) loc = loc * A.base + (int)((s >> td[2 + q]) & fm);
    const int *rp = T.rowptr + T.mrow_off[m];
    for (int e = rp[loc]; e < rp[loc + 1]; e++) n += T.col[e] != loc;
  }
  cnt[i] = n;
}

static std::mutex g_gtile_mu;  // first-time setup may be reached from several threads

// Chunks of whole rows with at most GT_CAP off-diagonal entries and
// GT_TPB / GT_Q rows each.  Not applicable: tensor bases (no table), more than
// 2^31 rows, entry ids beyond 16 bits, a row with more than GT_CAP
// entries.  Used by default for every basis that needs a canonical form, where
// it beat the row kernel at every size measured (ComplexF64, row -> tile):
// 5 x 5 torus at k = (1, 2), D = 208 012: 0.88 -> 0.39 ms; 6 x 5 torus,
// D = 5.17 M: 13.5 -> 13.1 ms.  ProjectedBasis keeps the row kernel: with no
// canonical form to compute, the row kernel's lanes look up the same term's
// images of consecutive rows, which share cache lines, and it is faster
// (J1-J2 ring, D = 10.4 M: 3.7 against 9.5 ms).  EDK_GENERIC_SCHEDULE=row or
// =tile overrides the choice.  Allocates and synchronises: run before capture.
static bool generic_setup_build(edk_op *op);

bool edk_generic_setup(edk_op *op) {
  std::lock_guard<std::mutex> lk(g_gtile_mu);
  if (op->gstate) return op->gstate > 0;
  // gstate stays 0 while the chunks are built, so a launch from another
  // thread waits on the mutex for the result; the single store below
  // publishes the chunks (1) or the row kernel (-1)
  const bool ok = generic_setup_build(op);
  op->gstate = ok ? 1 : -1;
  return ok;
}

static bool generic_setup_build(edk_op *op) {
  const edk_basis *b = op->basis;
  const char *env = getenv("EDK_GENERIC_SCHEDULE");
  if (env && strcmp(env, "row") == 0) return false;
  const bool force = env && strcmp(env, "tile") == 0;
  if (!force && b->kind == EDK_PROJECTED) return false;
  // (the per-row counts are copied to the host: at most 2^31 rows, 8 GB)
  if (b->kind == EDK_TENSOR || b->D == 0 || b->D > (1ll << 31) || op->terms.nnz >= GT_ABSENT) return false;
  edk_ctx *ctx = op->ctx;
  uint32_t *d_cnt = nullptr;
  std::vector<uint32_t> cnt(b->D);
  if (hipMalloc((void **)&d_cnt, b->D * sizeof(uint32_t)) != hipSuccess) return false;
  MatvecArgs a = op->args();
  k_row_entries<<<(unsigned)((b->D + 255) / 256), 256, 0, ctx->stream>>>(a, d_cnt);
  const bool okc = hipGetLastError() == hipSuccess &&
                   hipMemcpyAsync(cnt.data(), d_cnt, b->D * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream) ==
                       hipSuccess &&
                   hipStreamSynchronize(ctx->stream) == hipSuccess;
  (void)hipFree(d_cnt);
  if (!okc) return false;
  std::vector<uint32_t> cs;
  cs.push_back(0);
  uint32_t acc = 0, rows = 0;
  for (int64_t i = 0; i < b->D; i++) {
    if (cnt[i] > GT_CAP) return false;
    if (rows == GT_TPB / GT_Q || acc + cnt[i] > GT_CAP) {
      cs.push_back((uint32_t)i);
      acc = 0;
      rows = 0;
    }
    acc += cnt[i];
    rows++;
  }
  cs.push_back((uint32_t)b->D);
  if (hipMalloc((void **)&op->d_gchunk, cs.size() * sizeof(uint32_t)) != hipSuccess) return false;
  if (hipMemcpyAsync(op->d_gchunk, cs.data(), cs.size() * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream) !=
          hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    (void)hipFree(op->d_gchunk);
    op->d_gchunk = nullptr;
    return false;
  }
  op->ngchunk = (uint32_t)cs.size() - 1;
  return true;
}

template <typename VT, typename ST>
static void launch_gtile_t(const edk_op *op, const MatvecArgs &a, const VT *x, VT *y, int64_t nb, int accumulate,
                           hipStream_t st) {
  if (nb == 1) {
    k_matvec_gtile<VT, ST, 0><<<op->ngchunk, GT_TPB, 0, st>>>(a, op->d_gchunk, x, y, accumulate, 1, 0);
    return;
  }
  for (int64_t c = 0; c < nb;) {
    const int64_t left = nb - c;
    if (left >= 8) {
      k_matvec_gtile<VT, ST, 8><<<op->ngchunk, GT_TPB, 0, st>>>(a, op->d_gchunk, x, y, accumulate, nb, c);
      c += 8;
    } else if (left >= 4) {
      k_matvec_gtile<VT, ST, 4><<<op->ngchunk, GT_TPB, 0, st>>>(a, op->d_gchunk, x, y, accumulate, nb, c);
      c += 4;
    } else if (left >= 2) {
      k_matvec_gtile<VT, ST, 2><<<op->ngchunk, GT_TPB, 0, st>>>(a, op->d_gchunk, x, y, accumulate, nb, c);
      c += 2;
    } else {
      k_matvec_gtile<VT, ST, 1><<<op->ngchunk, GT_TPB, 0, st>>>(a, op->d_gchunk, x, y, accumulate, nb, c);
      c += 1;
    }
  }
}

// the tile form when it is set up (or can be: never under stream capture)
static bool launch_gtile(edk_op *op, int dtype, void *y, const void *x, int64_t nb, int accumulate,
                         hipStream_t st) {
  if (op->gstate == 0) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
    if (!edk_generic_setup(op)) return false;
  }
  if (op->gstate < 0) return false;
  const MatvecArgs a = op->args();
  if (dtype == EDK_C64)
    launch_gtile_t<double2, double>(op, a, (const double2 *)x, (double2 *)y, nb, accumulate, st);
  else
    launch_gtile_t<float2, float>(op, a, (const float2 *)x, (float2 *)y, nb, accumulate, st);
  return true;
}

// count of off-diagonal targets present in the sector (the NNZ_off figure)
__global__ void __launch_bounds__(256) k_count_offdiag(MatvecArgs A, unsigned long long *count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n = 0;
  if (i < A.D) {
    const TermsDev &T = A.T;
    uint64_t fm = (1ull << A.w) - 1;
    uint64_t s = A.kind == EDK_TENSOR ? edk_content_to_packed_d((uint64_t)i, A.L, A.base, A.w) : A.S.keys[i];
    for (int t = 0; t < T.nterms; t++) {
      const int *td = T.tdesc + t * EDK_TDESC;
      int ns = td[0], m = td[1];
      int loc = 0;
      for (int q = 0; q < ns; q++) loc = loc * A.base + (int)((s >> td[2 + q]) & fm);
      const int *rp = T.rowptr + T.mrow_off[m];
      for (int e = rp[loc]; e < rp[loc + 1]; e++) {
        int c = T.col[e];
        if (c == loc) continue;
        uint64_t s2 = s;
        int cc = c;
        for (int q = ns - 1; q >= 0; q--) {
          uint64_t nd = (uint64_t)(cc % A.base);
          cc /= A.base;
          int sh = td[2 + q];
          s2 = (s2 & ~(fm << sh)) | (nd << sh);
        }
        if (A.kind == EDK_TENSOR) n++;
        else if (A.kind == EDK_PROJECTED) n += edk_search(A.S, s2) >= 0;
        else {
          CanonOut o = edk_canon(A.C, s2);
          n += edk_search(A.S, o.rep) >= 0;
        }
      }
    }
  }
  // wave reduction then one atomic per wave (integer: order-independent)
  for (int o = 32; o > 0; o >>= 1) n += __shfl_down(n, o, 64);
  if ((threadIdx.x & 63) == 0 && n) atomicAdd(count, n);
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
MatvecArgs edk_op::args() const {
  MatvecArgs a;
  a.C = basis->canon();
  a.S = basis->search();
  a.T = terms;
  a.R = basis->d_R;
  a.kind = basis->kind;
  a.D = basis->D;
  a.L = basis->L;
  a.base = basis->base;
  a.w = basis->w;
  a.orbit_norms = basis->uploaded ? 0 : 1;
  return a;
}

extern "C" int edk_operator_create(edk_ctx *ctx, edk_basis *b, int64_t nterms, const int64_t *site_off,
                                   const int64_t *sites, const int64_t *colptr_off, const int64_t *colptr,
                                   const int64_t *nz_off, const int64_t *rowval, const double *nzval,
                                   edk_op **out) {
  if (!ctx || !b || !out || nterms < 0) return edk_fail(EDK_ERR_ARG, "edk_operator_create: bad argument");
  if (nterms > 0 && (!site_off || !sites || !colptr_off || !colptr || !nz_off))
    return edk_fail(EDK_ERR_ARG, "edk_operator_create: null term arrays");
  int L = b->L, base = b->base, w = b->w;
  // deduplicate local matrices (trans_inv_operator repeats one matrix L times)
  struct Mat {
    int n;
    std::vector<int64_t> cp, rv;
    std::vector<double> nz;
  };
  std::vector<Mat> mats;
  std::vector<int> tdesc(std::max<int64_t>(nterms, 1) * EDK_TDESC, 0);
  std::vector<uint64_t> tpw(std::max<int64_t>(nterms, 1) * EDK_MAXSITES, 0);
  bool all_real = true;
  for (int64_t t = 0; t < nterms; t++) {
    int64_t ns = site_off[t + 1] - site_off[t];
    if (ns < 1 || ns > EDK_MAXSITES) return edk_fail(EDK_ERR_UNSUPPORTED, "local terms act on 1..8 sites");
    int64_t n = 1;
    for (int q = 0; q < ns; q++) n *= base;
    if (n > 4096) return edk_fail(EDK_ERR_UNSUPPORTED, "local term dimension above 4096");
    Mat M;
    M.n = (int)n;
    M.cp.assign(colptr + colptr_off[t], colptr + colptr_off[t] + n + 1);
    int64_t nnz = nz_off[t + 1] - nz_off[t];
    if (M.cp[0] != 1 || M.cp[n] - 1 != nnz) return edk_fail(EDK_ERR_ARG, "inconsistent colptr");
    M.rv.assign(rowval + nz_off[t], rowval + nz_off[t] + nnz);
    M.nz.assign(nzval + 2 * nz_off[t], nzval + 2 * (nz_off[t] + nnz));
    for (int64_t q = 0; q < nnz; q++) {
      if (M.rv[q] < 1 || M.rv[q] > n) return edk_fail(EDK_ERR_ARG, "rowval out of range");
      if (M.nz[2 * q + 1] != 0.0) all_real = false;
    }
    int id = -1;
    for (size_t m = 0; m < mats.size() && id < 0; m++)
      if (mats[m].n == M.n && mats[m].cp == M.cp && mats[m].rv == M.rv && mats[m].nz == M.nz) id = (int)m;
    if (id < 0) {
      id = (int)mats.size();
      mats.push_back(M);
    }
    tdesc[t * EDK_TDESC + 0] = (int)ns;
    tdesc[t * EDK_TDESC + 1] = id;
    for (int q = 0; q < ns; q++) {
      int64_t site = sites[site_off[t] + q];
      if (site < 1 || site > L) return edk_fail(EDK_ERR_ARG, "site index out of 1:L");
      tdesc[t * EDK_TDESC + 2 + q] = (int)((L - site) * w);
      tpw[t * EDK_MAXSITES + q] = b->pw[L - site];
    }
  }
  // CSR of every distinct matrix: row r lists (c, M[r, c])
  std::vector<int> mrow_off, rowptr, col;
  std::vector<double2> val;
  for (auto &M : mats) {
    mrow_off.push_back((int)rowptr.size());
    std::vector<std::vector<std::pair<int, double2>>> rows(M.n);
    for (int c = 0; c < M.n; c++)
      for (int64_t q = M.cp[c] - 1; q < M.cp[c + 1] - 1; q++)
        rows[M.rv[q] - 1].push_back({c, make_double2(M.nz[2 * q], M.nz[2 * q + 1])});
    for (int r = 0; r < M.n; r++) {
      rowptr.push_back((int)col.size());
      for (auto &e : rows[r]) {
        col.push_back(e.first);
        val.push_back(e.second);
      }
    }
    rowptr.push_back((int)col.size());
  }
  if (mats.empty()) {
    mrow_off.push_back(0);
    rowptr.push_back(0);
  }
  if (col.empty()) {
    col.push_back(0);
    val.push_back(make_double2(0, 0));
  }
  edk_op *op = new edk_op();
  op->ctx = ctx;
  op->basis = b;
  op->all_real = all_real;
  op->nterms = (int)nterms;
  // EDK_GENERIC_MATVEC=1 forces the generic kernel (lets tests compare kernels)
  const char *gen = getenv("EDK_GENERIC_MATVEC");
  op->disable_fast = gen && gen[0] == '1';
  op->h_tdesc = tdesc;
  op->h_mrow_off = mrow_off;
  op->h_rowptr = rowptr;
  op->h_col = col;
  op->h_val = val;
  int rc = EDK_OK;
  EDK_HIP(hipSetDevice(ctx->device));
  auto up = [&](const void *h, size_t bytes, void **d) -> int {
    if (hipMalloc(d, std::max<size_t>(bytes, 16)) != hipSuccess) return edk_fail(EDK_ERR_OOM, "operator tables");
    if (bytes && hipMemcpyAsync(*d, h, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
      return edk_fail(EDK_ERR_HIP, "operator upload");
    return EDK_OK;
  };
  if (!rc) rc = up(tdesc.data(), tdesc.size() * sizeof(int), (void **)&op->d_tdesc);
  if (!rc) rc = up(tpw.data(), tpw.size() * sizeof(uint64_t), (void **)&op->d_tpw);
  if (!rc) rc = up(mrow_off.data(), mrow_off.size() * sizeof(int), (void **)&op->d_mrow_off);
  if (!rc) rc = up(rowptr.data(), rowptr.size() * sizeof(int), (void **)&op->d_rowptr);
  if (!rc) rc = up(col.data(), col.size() * sizeof(int), (void **)&op->d_col);
  if (!rc) rc = up(val.data(), val.size() * sizeof(double2), (void **)&op->d_val);
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = edk_fail(EDK_ERR_HIP, "operator upload sync");
  if (rc) {
    edk_operator_destroy(op);
    return rc;
  }
  op->terms.nterms = (int)nterms;
  op->terms.tdesc = op->d_tdesc;
  op->terms.tpw = op->d_tpw;
  op->terms.mrow_off = op->d_mrow_off;
  op->terms.rowptr = op->d_rowptr;
  op->terms.col = op->d_col;
  op->terms.val = op->d_val;
  op->terms.nmo = (int)mrow_off.size();
  op->terms.nrp = (int)rowptr.size();
  op->terms.nnz = (int)col.size();
  b->refcount++;
  *out = op;
  return EDK_OK;
}

extern "C" int edk_operator_destroy(edk_op *op) {
  if (!op) return EDK_OK;
  hipSetDevice(op->ctx->device);
  hipStreamSynchronize(op->ctx->stream);
  hipFree(op->d_tdesc);
  hipFree(op->d_tpw);
  hipFree(op->d_mrow_off);
  hipFree(op->d_rowptr);
  hipFree(op->d_col);
  hipFree(op->d_val);
  hipFree(op->d_scratch);
  hipFree(op->d_gchunk);
  edk_fast_release(op);
  edk_perm_release(op);
  edk_csr_release(op);
  if (op->basis) op->basis->refcount--;
  delete op;
  return EDK_OK;
}

extern "C" int edk_operator_count_offdiag(edk_op *op, int64_t *nnz) {
  if (!op || !nnz) return edk_fail(EDK_ERR_ARG, "null argument");
  edk_ctx *ctx = op->ctx;
  EDK_HIP(hipSetDevice(ctx->device));
  unsigned long long *d = nullptr, h = 0;
  EDK_HIP(hipMalloc((void **)&d, 8));
  EDK_HIP(hipMemsetAsync(d, 0, 8, ctx->stream));
  MatvecArgs a = op->args();
  if (a.D > 0) k_count_offdiag<<<(unsigned)((a.D + 255) / 256), 256, 0, ctx->stream>>>(a, d);
  EDK_HIP(hipGetLastError());
  EDK_HIP(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, ctx->stream));
  EDK_HIP(hipStreamSynchronize(ctx->stream));
  hipFree(d);
  *nnz = (int64_t)h;
  return EDK_OK;
}

static thread_local edk_mv_hook *tl_mv_hook = nullptr;
void edk_set_mv_hook(edk_mv_hook *h) { tl_mv_hook = h; }

int edk_matvec_launch(edk_op *op, int dtype, void *y, const void *x, int64_t nb, int accumulate, hipStream_t st) {
  MatvecArgs a = op->args();
  if (a.D == 0) return EDK_OK;
  if (tl_mv_hook && nb == 1 && !accumulate) {
    const int rc = tl_mv_hook->launch(op, dtype, y, x, st);
    if (rc >= 0) return rc;
  }
  // several right-hand sides with the assembled matrix cached (sparse!)
  if (nb > 1 && edk_matvec_csr(op, dtype, y, x, nb, accumulate, st)) return EDK_OK;
  if (edk_matvec_fast(op, dtype, y, x, nb, accumulate, st)) return EDK_OK;
  if (nb == 1 && edk_matvec_perm(op, dtype, y, x, accumulate, st)) {
    EDK_HIP(hipGetLastError());
    return EDK_OK;
  }
  if (launch_gtile(op, dtype, y, x, nb, accumulate, st)) {
    EDK_HIP(hipGetLastError());
    return EDK_OK;
  }
  if (dtype == EDK_C64)
    launch_generic<double2, double>(a, (const double2 *)x, (double2 *)y, nb, accumulate, st);
  else
    launch_generic<float2, float>(a, (const float2 *)x, (float2 *)y, nb, accumulate, st);
  EDK_HIP(hipGetLastError());
  return EDK_OK;
}

extern "C" int edk_matvec(edk_op *op, int dtype, void *y, const void *x, int64_t nb, int accumulate) {
  if (!op || !y || !x || nb < 1) return edk_fail(EDK_ERR_ARG, "edk_matvec: bad argument");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "edk_matvec: dtype must be EDK_C64 or EDK_C32");
  if (x == y) return edk_fail(EDK_ERR_ARG, "edk_matvec: x and y must not alias");
  EDK_HIP(hipSetDevice(op->ctx->device));
  return edk_matvec_launch(op, dtype, y, x, nb, accumulate);
}

extern "C" int edk_matvec_multi(edk_op *const *ops, int n, int dtype, void *const *ys, const void *const *xs) {
  if (!ops || !ys || !xs || n < 1 || n > 8) return edk_fail(EDK_ERR_ARG, "edk_matvec_multi: 1..8 operators");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "edk_matvec_multi: bad dtype");
  for (int c = 0; c < n; c++) {
    if (!ops[c] || !ys[c] || !xs[c]) return edk_fail(EDK_ERR_ARG, "edk_matvec_multi: null argument");
    if (ops[c]->basis != ops[0]->basis) return edk_fail(EDK_ERR_ARG, "edk_matvec_multi: operators on different bases");
    for (int q = 0; q < n; q++)
      if (xs[q] == ys[c]) return edk_fail(EDK_ERR_ARG, "edk_matvec_multi: an input aliases an output");
  }
  edk_ctx *ctx = ops[0]->ctx;
  EDK_HIP(hipSetDevice(ctx->device));
  const int64_t D = ops[0]->basis->D;
  if (D == 0) return EDK_OK;
  void *X = nullptr;
  const size_t bytes = (size_t)D * 8 * (dtype == EDK_C64 ? 16 : 8);
  if (edk_dalloc(ctx, &X, bytes) != hipSuccess) return edk_fail(EDK_ERR_OOM, "edk_matvec_multi block");
  int rc = edk_fast_matvec_mc(ops, n, dtype, xs, ys, X, ctx->stream);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc) rc = edk_fail(EDK_ERR_HIP, "edk_matvec_multi");
  edk_dfree(ctx, X);
  return rc;
}

extern "C" int edk_matvec_host(edk_op *op, int dtype, void *yh, const void *xh, int64_t nb, int accumulate) {
  if (!op || !yh || !xh || nb < 1) return edk_fail(EDK_ERR_ARG, "edk_matvec_host: bad argument");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "edk_matvec_host: bad dtype");
  edk_ctx *ctx = op->ctx;
  size_t bytes = (size_t)op->basis->D * nb * (dtype == EDK_C64 ? 16 : 8);
  void *dx = nullptr, *dy = nullptr;
  EDK_HIP(hipSetDevice(ctx->device));
  if (hipMalloc(&dx, std::max<size_t>(bytes, 16)) != hipSuccess) return edk_fail(EDK_ERR_OOM, "matvec_host x");
  if (hipMalloc(&dy, std::max<size_t>(bytes, 16)) != hipSuccess) {
    hipFree(dx);
    return edk_fail(EDK_ERR_OOM, "matvec_host y");
  }
  int rc = EDK_OK;
  if (hipMemcpyAsync(dx, xh, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      hipMemcpyAsync(dy, yh, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
    rc = edk_fail(EDK_ERR_HIP, "matvec_host copy in");
  if (!rc) rc = edk_matvec_launch(op, dtype, dy, dx, nb, accumulate);
  if (!rc && hipMemcpyAsync(yh, dy, bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
    rc = edk_fail(EDK_ERR_HIP, "matvec_host copy out");
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = edk_fail(EDK_ERR_HIP, "matvec_host sync");
  hipFree(dx);
  hipFree(dy);
  return rc;
}
