// edk_fast.hip -- specialised gather kernels for one vector, and the host
// setup they share with the block kernels (edk_fast.h describes the method).
#include <mutex>

#include "edk_fast.h"

template <int KIND, int PH, int NB, int LT, typename VT, typename ST>
__global__ void __launch_bounds__(FAST_TPB, NB >= 8 ? 4 : (NB >= 4 ? 7 : 8))
    k_matvec_fast(FastArgs A, const VT *__restrict__ x, VT *__restrict__ y, int accumulate) {
  constexpr bool CPLX = (PH == 2);
  // orbit norms and phases staged in LDS (read once per target)
  __shared__ double sR[2 * 64];
  __shared__ double2 sP[PH != 0 ? 64 : 1];
  if (KIND != EDK_PROJECTED) {
    int nr = (KIND == EDK_TRANSLATION_PARITY ? 2 : 1) * A.len;
    if ((int)threadIdx.x < nr) sR[threadIdx.x] = A.Rtab[threadIdx.x];
    if (PH != 0 && KIND == EDK_TRANSLATIONAL && (int)threadIdx.x < A.len) sP[threadIdx.x] = A.ph[threadIdx.x];
    __syncthreads();
  }
  FastArgs B = A;
  B.Rtab = sR;
  if (PH != 0 && KIND == EDK_TRANSLATIONAL) B.ph = sP;
  uint32_t i = blockIdx.x * FAST_TPB + threadIdx.x;
  if (i >= A.D) return;
  const FastTab &T = *A.tab;
  const uint32_t s = A.keys[i];
  const uint32_t ns = ~s & A.mask;
  ST dg = 0;
  ST orr = 0, oi = 0;
  // families and entries are walked by uniform loops (scalar table reads);
  // only the bit loop over one entry's positions diverges between lanes
  for (int f = 0; f < T.nfam; f++) {
    const int d = T.fam[f].d;
    const uint32_t r = rotr32(s, A.L - d, A.L, A.mask);  // rotl by d
    const uint32_t nr = ~r & A.mask;
    dg += (ST)(T.fam[f].diag[0] * __popc(ns & nr) + T.fam[f].diag[1] * __popc(ns & r) +
               T.fam[f].diag[2] * __popc(s & nr) + T.fam[f].diag[3] * __popc(s & r));
    for (int e = 0; e < T.fam[f].nent; e++) {
      const int flip = T.fam[f].ent[e].flip;
      const double val = T.fam[f].ent[e].val;
      uint32_t bits = ent_mask(T.fam[f].ent[e].locs, s, ns, r, A.mask);
      while (bits) {
        uint32_t rep[NB], lo[NB], hi[NB];
        bool ok[NB];
        ST cfr[NB], cfi[CPLX ? NB : 1];
#pragma unroll
        for (int q = 0; q < NB; q++) {
          ok[q] = bits != 0;
          int p = __ffs(bits) - 1;
          bits &= bits - 1;
          int p2 = p - d;
          if (LT == 32) p2 &= 31;
          else if (p2 < 0) p2 += A.L;
          uint32_t t = s;
          if (flip & 2) t ^= 1u << p;
          if (flip & 1) t ^= 1u << p2;
          double cr = 0.0, ci = 0.0;
          rep[q] = 0;
          if (ok[q]) {
            if (LT == 32) ok[q] = fast_canon32<KIND, PH>(B, t, rep[q], cr, ci);
            else ok[q] = fast_canon<KIND, PH>(B, t, rep[q], cr, ci);
          }
          lo[q] = 0;
          hi[q] = 0;
          if (ok[q] && A.region) {
            region_bounds(A, rep[q], lo[q], hi[q]);
          } else if (ok[q]) {
            uint32_t bkt = rep[q] >> A.dshift;
            if (bkt < A.ndir) {
              lo[q] = A.dir[bkt];
              hi[q] = A.dir[bkt + 1];
            }
          }
          cfr[q] = (ST)(val * cr);
          if (CPLX) cfi[q] = (ST)(val * ci);
        }
        uint32_t jj[NB];
        if (A.region) {
#pragma unroll
          for (int q = 0; q < NB; q++) jj[q] = region_find(A, rep[q], lo[q], hi[q]);
        } else {
#pragma unroll
          for (int q = 0; q < NB; q++) jj[q] = (uint32_t)fast_find(A, rep[q], lo[q], hi[q]);
        }
#pragma unroll
        for (int q = 0; q < NB; q++) {
          if (jj[q] != 0xffffffffu) {
            VT xj = x[jj[q]];
            if (CPLX) {
              orr += cfr[q] * (ST)xj.x - cfi[q] * (ST)xj.y;
              oi += cfr[q] * (ST)xj.y + cfi[q] * (ST)xj.x;
            } else {
              orr += cfr[q] * (ST)xj.x;
              oi += cfr[q] * (ST)xj.y;
            }
          }
        }
      }
    }
  }
  VT xi = x[i];
  ST invR = (KIND == EDK_PROJECTED) ? (ST)1 : (ST)(1.0 / A.Rrow[i]);
  ST yr = dg * (ST)xi.x + orr * invR;
  ST yi = dg * (ST)xi.y + oi * invR;
  VT out;
  if (accumulate) {
    VT y0 = y[i];
    out.x = y0.x + yr;
    out.y = y0.y + yi;
  } else {
    out.x = yr;
    out.y = yi;
  }
  y[i] = out;
}

template <int KIND, int PH, int LT, typename VT, typename ST>
__global__ void __launch_bounds__(BAL_TPB, ((LT == 32 || LT == 1) && PH == 0 && KIND != EDK_PROJECTED) ? 7 : (((LT == 32 || LT == 1) && PH == 2) ? 6 : 1)) k_matvec_bal(FastArgs A, const VT *__restrict__ x, VT *__restrict__ y,
                                                         int accumulate) {
  constexpr bool CPLX = (PH == 2);
  __shared__ double sR[2 * 64];
  __shared__ double2 sP[PH != 0 ? 64 : 1];
  __shared__ double sV[FAST_MAXFAM * FAST_MAXENT];
  __shared__ uint32_t tgt[BAL_CAP];
  __shared__ uint8_t ent[BAL_CAP];
  __shared__ VT contr[BAL_CAP];  // contributions, (re, im) together
  __shared__ uint32_t wsum[BAL_TPB / 64];
  __shared__ uint32_t ndef;
  // L = 32, phase mode 0: run-length candidates, with the few targets that
  // have more than two longest zero runs deferred to a block-wide full search.
  // Results go back to LDS: the representative over the target and the
  // coefficient into the imaginary half of the target's own contr slot (only
  // its owner reads it, then overwrites it with the contribution); the
  // deferred queue lives in the real halves, dead before the gathers write
  // (LT == 32: L = 32; LT == 1: 16 <= L < 32; both with unit cell 1)
  constexpr bool RUNS = (LT == 32 || LT == 1) && KIND != EDK_PROJECTED;
  constexpr int DQS = sizeof(VT) / sizeof(uint32_t);  // queue entry d at the real half of contr[d]
  uint32_t *dq = reinterpret_cast<uint32_t *>(contr);
  // complex coefficients fill their contr slot: the queue then has its own array
  __shared__ uint16_t dqc[(RUNS && CPLX) ? BAL_CAP : 1];
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
  const uint32_t cb = bal_chunk(blockIdx.x, A.nchunk, A.xcdg);
  const uint32_t r0 = A.chunk[cb], nrow = A.chunk[cb + 1] - r0;
  // phase 1: row threads count and write their targets; the row's own x_i
  // and 1/R_i are loaded now so their latency hides behind phases 1 and 2
  uint32_t s = 0, n = 0;
  const bool own = (uint32_t)tid < nrow;
  VT xi{};
  ST invR = 1;
  if (own) {
    xi = x[r0 + tid];
    if (KIND != EDK_PROJECTED) invR = (ST)(1.0 / A.Rrow[r0 + tid]);
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
  }
  __syncthreads();
  // phase 2: BAL_K targets per thread, no divergence
  uint32_t rep[BAL_K], jj[BAL_K];
  bool ok[BAL_K];
  ST cfr[BAL_K], cfi[CPLX ? BAL_K : 1];
  if constexpr (RUNS) {
    // one target at a time, results straight to LDS: little live state.  A
    // representative outside the sector is marked by ent = 0xff.
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
          if (CPLX) {
            VT c;
            c.x = (ST)(v * cr);
            c.y = (ST)(v * ci);
            contr[j] = c;
          } else {
            contr[j].y = (ST)(v * cr);
          }
        } else if (st == 0) {
          ent[j] = 0xff;
        } else {
          const uint32_t d = atomicAdd(&ndef, 1u);
          if (CPLX) dqc[d] = (uint16_t)j;
          else dq[DQS * d] = j;
        }
      }
    }
    __syncthreads();
    const uint32_t nd = ndef;
    for (uint32_t d = tid; d < nd; d += BAL_TPB) {
      const uint32_t j = CPLX ? (uint32_t)dqc[d] : dq[DQS * d];
      uint32_t r = 0;
      double cr = 0.0, ci = 0.0;
      const bool okb = (LT == 32) ? fast_canon32<KIND, PH>(B, tgt[j], r, cr, ci)
                                  : fast_canon_w<KIND, PH>(B, tgt[j], r, cr, ci);
      if (okb) {
        tgt[j] = r;
        const double v = sV[ent[j]];
        if (CPLX) {
          VT c;
          c.x = (ST)(v * cr);
          c.y = (ST)(v * ci);
          contr[j] = c;
        } else {
          contr[j].y = (ST)(v * cr);
        }
      } else {
        ent[j] = 0xff;
      }
    }
    // results visible; the queue is dead before the contributions reuse it
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BAL_K; q++) {
      const uint32_t j = tid + BAL_TPB * q;
      ok[q] = j < total && ent[j] != 0xff;
      rep[q] = ok[q] ? tgt[j] : 0u;
      if (CPLX) {
        const VT c = contr[j];
        cfr[q] = ok[q] ? c.x : (ST)0;
        cfi[q] = ok[q] ? c.y : (ST)0;
      } else {
        cfr[q] = ok[q] ? contr[j].y : (ST)0;
      }
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
  // gathers issued together; an absent target reads x[0] and drops it by
  // selection, so a non-finite x[0] cannot leak into the sum
  VT xg[BAL_K];
#pragma unroll
  for (int q = 0; q < BAL_K; q++) xg[q] = x[jj[q] != HASH_EMPTY ? jj[q] : 0u];
#pragma unroll
  for (int q = 0; q < BAL_K; q++) {
    const uint32_t j = tid + BAL_TPB * q;
    ST cr_ = 0, ci_ = 0;
    const VT xj = xg[q];
    if (CPLX) {
      cr_ = cfr[q] * (ST)xj.x - cfi[q] * (ST)xj.y;
      ci_ = cfr[q] * (ST)xj.y + cfi[q] * (ST)xj.x;
    } else {
      cr_ = cfr[q] * (ST)xj.x;
      ci_ = cfr[q] * (ST)xj.y;
    }
    if (jj[q] == HASH_EMPTY) {
      cr_ = 0;
      ci_ = 0;
    }
    if (j < total) {
      VT cv;
      cv.x = cr_;
      cv.y = ci_;
      contr[j] = cv;
    }
  }
  __syncthreads();
  // phase 3: row sums in generator order, diagonal, output
  if (own) {
    const uint32_t i = r0 + tid;
    const uint32_t ns = ~s & A.mask;
    ST dg = 0;
    for (int f = 0; f < T.nfam; f++) {
      uint32_t r = rotr_lt<LT>(s, A.L - T.fam[f].d, A.L, A.mask);
      uint32_t nr = ~r & A.mask;
      dg += (ST)(T.fam[f].diag[0] * __popc(ns & nr) + T.fam[f].diag[1] * __popc(ns & r) +
                 T.fam[f].diag[2] * __popc(s & nr) + T.fam[f].diag[3] * __popc(s & r));
    }
    ST orr = 0, oi = 0;
    // in generator order, four reads in flight at a time
    uint32_t q = off;
    const uint32_t qe = off + n;
    for (; q + 4 <= qe; q += 4) {
      const VT c0 = contr[q], c1 = contr[q + 1], c2 = contr[q + 2], c3 = contr[q + 3];
      orr += c0.x;
      oi += c0.y;
      orr += c1.x;
      oi += c1.y;
      orr += c2.x;
      oi += c2.y;
      orr += c3.x;
      oi += c3.y;
    }
    for (; q < qe; q++) {
      const VT cv = contr[q];
      orr += cv.x;
      oi += cv.y;
    }
    ST yr = dg * (ST)xi.x + orr * invR;
    ST yi = dg * (ST)xi.y + oi * invR;
    VT out;
    if (accumulate) {
      VT y0 = y[i];
      out.x = y0.x + yr;
      out.y = y0.y + yi;
    } else {
      out.x = yr;
      out.y = yi;
    }
    y[i] = out;
  }
}

// number of off-diagonal targets the generator produces for each row
template <typename CT>
__global__ void k_row_targets(FastArgs A, CT *cnt) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.D) return;
  const FastTab &T = *A.tab;
  uint32_t s = A.keys[i], ns = ~s & A.mask;
  int n = 0;
  for (int f = 0; f < T.nfam; f++) {
    uint32_t r = rotr32(s, A.L - T.fam[f].d, A.L, A.mask);
    for (int e = 0; e < T.fam[f].nent; e++) n += __popc(ent_mask(T.fam[f].ent[e].locs, s, ns, r, A.mask));
  }
  cnt[i] = (CT)min(n, (int)(CT)~(CT)0);
}

__global__ void k_keys32(const uint64_t *k64, int64_t D, uint32_t *k32) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < D) k32[i] = (uint32_t)k64[i];
}

__global__ void k_dir32(const uint32_t *keys, int64_t D, int dshift, uint32_t ndir, uint32_t *dir) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > (int64_t)ndir) return;
  if (b == (int64_t)ndir) {
    dir[b] = (uint32_t)D;
    return;
  }
  uint64_t key = (uint64_t)b << dshift;
  int64_t lo = 0, hi = D;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((uint64_t)keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  dir[b] = (uint32_t)lo;
}

// colex ranks of the keys (sorted, so the ranks increase)
__global__ void k_rank_keys(const uint32_t *keys, uint32_t D, const uint32_t *T, uint32_t dmask, int rlo, int ng,
                            uint32_t *rk) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < D) rk[i] = rank_full(T, (keys[i] & dmask) >> rlo, ng);
}

// granule g: representatives of rank below 64 g, membership of 64 g .. 64 g + 63
// (bitmap words 0 and 1; the slot's last word stays 0)
__global__ void k_gran_build(const uint32_t *rk, uint32_t D, uint32_t ngran, uint4 *gran) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngran) return;
  const uint64_t r0 = (uint64_t)g * RK_GRAN;
  uint32_t lo = 0, hi = D;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if ((uint64_t)rk[mid] < r0) lo = mid + 1;
    else hi = mid;
  }
  uint32_t w[3] = {0u, 0u, 0u};
  for (uint32_t i = lo; i < D && (uint64_t)rk[i] < r0 + RK_GRAN; i++) {
    const uint32_t o = (uint32_t)(rk[i] - r0);
    w[o >> 5] |= 1u << (o & 31u);
  }
  gran[g] = make_uint4(lo, w[0], w[1], w[2]);
}

static FastState *fast_state(edk_op *op) {
  if (!op->fast) op->fast = new FastState();
  return (FastState *)op->fast;
}

void edk_fast_release(edk_op *op) {
  FastState *F = (FastState *)op->fast;
  if (!F) return;
  (void)hipFree(F->d_tab);
  (void)hipFree(F->d_coltab);
  (void)hipFree(F->d_keys32);
  (void)hipFree(F->d_dir);
  (void)hipFree(F->d_region);
  (void)hipFree(F->d_fine);
  (void)hipFree(F->d_chunk);
  (void)hipFree(F->d_gran);
  (void)hipFree(F->d_rtab);
  delete F;
  op->fast = nullptr;
}

// Decompose the operator into translation-invariant two-site families.
static bool find_families(const edk_op *op, FastTab &tab) {
  const edk_basis *b = op->basis;
  int L = b->L;
  if (op->nterms == 0 || op->nterms % L) return false;
  struct Key {
    int mat, d;
  };
  std::vector<Key> keys;
  std::vector<std::vector<int>> starts;
  for (int t = 0; t < op->nterms; t++) {
    const int *td = &op->h_tdesc[t * EDK_TDESC];
    if (td[0] != 2) return false;
    int s1 = L - td[2], s2 = L - td[3];  // sites from field shifts (w == 1)
    int d = ((s2 - s1) % L + L) % L;
    if (d == 0) return false;
    int id = -1;
    for (size_t k = 0; k < keys.size(); k++)
      if (keys[k].mat == td[1] && keys[k].d == d) id = (int)k;
    if (id < 0) {
      id = (int)keys.size();
      keys.push_back({td[1], d});
      starts.push_back({});
    }
    starts[id].push_back(s1);
  }
  if ((int)keys.size() > FAST_MAXFAM) return false;
  tab.nfam = (int)keys.size();
  for (size_t k = 0; k < keys.size(); k++) {
    std::vector<int> &st = starts[k];
    if ((int)st.size() != L) return false;
    std::sort(st.begin(), st.end());
    for (int i = 0; i < L; i++)
      if (st[i] != i + 1) return false;
    FastFam &F = tab.fam[k];
    F.d = keys[k].d;
    F.nent = 0;
    for (int a = 0; a < 4; a++) F.diag[a] = 0.0;
    int m = keys[k].mat;
    const int *rp = &op->h_rowptr[op->h_mrow_off[m]];
    for (int loc = 0; loc < 4; loc++) {
      for (int q = rp[loc]; q < rp[loc + 1]; q++) {
        int c = op->h_col[q];
        double2 v = op->h_val[q];
        if (v.y != 0.0) return false;
        if (c == loc) {
          F.diag[loc] += v.x;
          continue;
        }
        // entries that flip the same sites with the same value share one mask
        int hit = -1;
        for (int q2 = 0; q2 < F.nent; q2++)
          if (F.ent[q2].flip == (loc ^ c) && F.ent[q2].val == v.x) hit = q2;
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
  }
  return true;
}

static std::mutex g_setup_mu;  // first-time setup may be reached from several threads

static bool fast_setup_build(edk_op *op);

// Launches read the published state without the lock; only a launch that
// finds setup not yet done takes the mutex (and waits for a build running in
// another thread).  The single store of fstate publishes the state.
static bool fast_setup(edk_op *op) {
  const int st = op->fstate.load(std::memory_order_acquire);
  if (st) return st > 0;
  std::lock_guard<std::mutex> setup_lock(g_setup_mu);
  if (op->fstate.load(std::memory_order_relaxed)) return op->fstate.load(std::memory_order_relaxed) > 0;
  const bool ok = fast_setup_build(op);
  op->fstate.store(ok ? 1 : -1, std::memory_order_release);
  return ok;
}

// Rank directory over sorted 32-bit representatives of L sites: applies
// when every representative has one popcount; the domain leaves out bit 0
// when every key has it set and the top bit when every key has it clear (true
// of smallest rotations).  Fills A's rank fields (A.gran stays null when the
// directory does not apply) and the two device arrays.  Returns false only on
// a device error.
bool edk_rank_build(edk_ctx *ctx, int L, const std::vector<uint32_t> &keys, const uint32_t *d_keys32, FastArgs &A,
                    uint4 **d_gran, uint32_t **d_rtab) {
  A.gran = nullptr;
  if (keys.empty() || L < 1 || L > 32) return true;
  const uint32_t top = 1u << (L - 1);
  const int pc = __builtin_popcount(keys[0]);
  bool all1 = true, all0 = true;
  for (uint32_t k : keys) {
    if (__builtin_popcount(k) != pc) return true;
    all1 = all1 && (k & 1u);
    all0 = all0 && !(k & top);
  }
  const int rlo = all1 ? 1 : 0, rbits = L - rlo - (all0 ? 1 : 0), rn = pc - rlo;
  if (rn < 0 || rn >= RK_NCNT || rbits < 1) return true;
  // Pascal's triangle to n = 35, then the 3-bit group table
  uint64_t Cb[36][36] = {};
  for (int n = 0; n < 36; n++) {
    Cb[n][0] = 1;
    for (int k = 1; k <= n; k++) Cb[n][k] = Cb[n - 1][k - 1] + (k < n ? Cb[n - 1][k] : 0);
  }
  if (Cb[rbits][rn] >= 0xffffffffull - RK_GRAN) return true;
  std::vector<uint32_t> T(RK_NT, 0u);
  for (int g = 0; g < RK_NGRP; g++)
    for (int c = 0; c < RK_NCNT; c++)
      for (int v = 0; v < 8; v++) {
        uint64_t t = 0;
        for (int b = 0, j = 0; b < 3; b++)
          if ((v >> b) & 1) {
            const int n = 3 * g + b, k = c + j + 1;
            t += (k <= n && n < 36) ? Cb[n][k] : 0;
            j++;
          }
        T[(g * RK_NCNT + c) * 8 + v] = (uint32_t)t;
      }
  A.rlo = rlo;
  A.rbits = rbits;
  A.rn = rn;
  A.rng = (rbits + 2) / 3;
  A.fix1 = all1 ? 1u : 0u;
  A.fix0 = all0 ? top : 0u;
  A.dmask = (uint32_t)((((uint64_t)1 << rbits) - 1) << rlo);
  const uint32_t D = (uint32_t)keys.size();
  uint32_t *d_rk = nullptr;
  if (hipMalloc((void **)d_rtab, RK_NT * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void **)&d_rk, (size_t)D * sizeof(uint32_t)) != hipSuccess) {
    (void)hipFree(d_rk);
    return false;
  }
  uint32_t last = 0;
  bool ok = hipMemcpyAsync(*d_rtab, T.data(), RK_NT * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream) ==
            hipSuccess;
  if (ok) {
    k_rank_keys<<<(D + 255) / 256, 256, 0, ctx->stream>>>(d_keys32, D, *d_rtab, A.dmask, rlo, A.rng, d_rk);
    ok = hipMemcpyAsync(&last, d_rk + (D - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
         hipStreamSynchronize(ctx->stream) == hipSuccess;
  }
  const uint32_t ngran = last / RK_GRAN + 1;
  ok = ok && hipMalloc((void **)d_gran, (size_t)ngran * sizeof(uint4)) == hipSuccess;
  if (ok) {
    k_gran_build<<<(ngran + 255) / 256, 256, 0, ctx->stream>>>(d_rk, D, ngran, *d_gran);
    ok = hipStreamSynchronize(ctx->stream) == hipSuccess;
  }
  (void)hipFree(d_rk);
  if (!ok) return false;
  A.gran = *d_gran;
  A.ngran = ngran;
  A.rtab = *d_rtab;
  return true;
}

// The directory for the run-length canonicalisers: unit cell 1, 16 <= L <=
// 32, translation kinds.
static bool rank_setup(edk_op *op, FastState *S, const std::vector<uint32_t> &keys) {
  const edk_basis *b = op->basis;
  S->args.gran = nullptr;
  if (b->A != 1 || b->L < 16 || b->L > 32 || (b->kind != EDK_TRANSLATION_PARITY && b->kind != EDK_TRANSLATIONAL))
    return true;
  return edk_rank_build(op->ctx, b->L, keys, S->d_keys32, S->args, &S->d_gran, &S->d_rtab);
}

static bool fast_setup_build(edk_op *op) {
  FastState *S = fast_state(op);
  const edk_basis *b = op->basis;
  if (b->base != 2 || b->L > 32 || b->uploaded) return false;
  if (b->kind != EDK_TRANSLATION_PARITY && b->kind != EDK_TRANSLATIONAL && b->kind != EDK_PROJECTED) return false;
  if (b->D == 0 || b->D >= (1ll << 32) - 1) return false;
  FastTab tab;
  if (!find_families(op, tab)) return false;
  FastArgs &A = S->args;
  A.kind = b->kind;
  A.L = b->L;
  A.A = b->A;
  A.len = b->len;
  A.kp = b->kp;
  A.P = b->P;
  A.mask = b->L == 32 ? 0xffffffffu : ((1u << b->L) - 1);
  A.ndiv = 0;
  for (int p = 1; p < b->len && A.ndiv < 16; p++)
    if (b->len % p == 0) A.divs[A.ndiv++] = p;
  // lane schedule: targets in flight per lane (the variants override it);
  // directory: about one slot per representative; lookups binary-search
  const int keys_per_bucket = 1;
  A.scan = 0;
  A.batch = 2;
  S->cplx = (b->kind == EDK_TRANSLATIONAL && !(b->kp == 0 || 2 * b->kp == b->len));
  edk_ctx *ctx = op->ctx;
  if (hipMalloc((void **)&S->d_tab, sizeof(FastTab)) != hipSuccess) return false;
  if (hipMemcpyAsync(S->d_tab, &tab, sizeof(FastTab), hipMemcpyHostToDevice, ctx->stream) != hipSuccess) return false;
  // this operator's values for the multi-coefficient kernel, and the hash of
  // its structure (operators with equal hashes on one basis share launches)
  {
    FastColTab ct{};
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
    mix((uint64_t)tab.nfam);
    for (int f = 0; f < tab.nfam; f++) {
      mix((uint64_t)tab.fam[f].d);
      mix((uint64_t)tab.fam[f].nent);
      for (int q = 0; q < 4; q++) ct.diag[f * 4 + q] = tab.fam[f].diag[q];
      for (int e = 0; e < tab.fam[f].nent; e++) {
        mix((uint64_t)tab.fam[f].ent[e].locs);
        mix((uint64_t)tab.fam[f].ent[e].flip);
        ct.val[f * FAST_MAXENT + e] = tab.fam[f].ent[e].val;
      }
    }
    S->signature = h;
    if (hipMalloc((void **)&S->d_coltab, sizeof(FastColTab)) != hipSuccess) return false;
    if (hipMemcpyAsync(S->d_coltab, &ct, sizeof(FastColTab), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
      return false;
  }
  // keys + 64 padding entries (the wide scan may read past the end)
  if (hipMalloc((void **)&S->d_keys32, (b->D + 64) * sizeof(uint32_t)) != hipSuccess) return false;
  if (hipMemsetAsync(S->d_keys32 + b->D, 0xff, 64 * sizeof(uint32_t), ctx->stream) != hipSuccess) return false;
  k_keys32<<<(unsigned)((b->D + 255) / 256), 256, 0, ctx->stream>>>(b->d_keys, b->D, S->d_keys32);
  // directory over the 32-bit keys: about one slot per keys_per_bucket representatives
  int64_t target = std::max<int64_t>(1024, b->D / std::max(1, keys_per_bucket));
  int sh = 0;
  while (sh < 31 && (int64_t)(b->maxkey >> sh) + 1 > target) sh++;
  uint32_t ndir = (uint32_t)((b->maxkey >> sh) + 1);
  if (hipMalloc((void **)&S->d_dir, ((size_t)ndir + 1) * sizeof(uint32_t)) != hipSuccess) return false;
  k_dir32<<<(unsigned)((ndir + 256) / 256), 256, 0, ctx->stream>>>(S->d_keys32, b->D, sh, ndir, S->d_dir);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return false;
  A.tab = S->d_tab;
  A.Rtab = b->d_Rtab;
  A.ph = b->d_ph;
  A.keys = S->d_keys32;
  A.dir = S->d_dir;
  A.dshift = sh;
  A.ndir = ndir;
  A.D = (uint32_t)b->D;
  A.Rrow = b->d_R;
  A.region = nullptr;
  A.fine = nullptr;
  A.gran = nullptr;
  {
    // two-level directory (4096 regions of 2^20 key values); fine slots
    // ~ one per representative; entries are 16-bit offsets from the region's
    // first index, so a region may hold at most 65535 representatives
    std::vector<uint32_t> keys(b->D);
    if (hipMemcpyAsync(keys.data(), S->d_keys32, b->D * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      return false;
    // regions: the fewest top bits (12..16) that keep every region under 65536 keys
    int L = b->L, RS = 0;
    std::vector<uint32_t> first;
    bool fits = false;
    for (int RB = 12; RB <= 16 && !fits; RB++) {
      RS = std::max(L - RB, 0);
      uint32_t nreg = 1u << std::min(RB, L);
      first.assign(nreg + 1, 0);
      for (int64_t i = 0, r = 0; r <= (int64_t)nreg; r++) {
        while (i < b->D && (int64_t)(keys[i] >> RS) < r) i++;
        first[r] = (uint32_t)i;
      }
      fits = true;
      for (uint32_t r = 0; r < nreg; r++)
        if (first[r + 1] - first[r] > 65535) fits = false;
    }
    if (fits) {
      uint32_t nreg = (uint32_t)first.size() - 1;
      std::vector<uint2> region(nreg);
      std::vector<uint16_t> fine;
      for (uint32_t r = 0; r < nreg; r++) {
        uint32_t n = first[r + 1] - first[r];
        int sh = RS;
        while (sh > 0 && (1u << (RS - sh)) < std::max<uint32_t>(n, 1)) sh--;
        uint32_t nf = 1u << (RS - sh);
        region[r] = make_uint2(first[r], ((uint32_t)fine.size() << 5) | (uint32_t)sh);
        uint32_t i = first[r];
        for (uint32_t fb = 0; fb <= nf; fb++) {
          uint64_t lim = ((uint64_t)r << RS) + ((uint64_t)fb << sh);
          while (i < first[r + 1] && (uint64_t)keys[i] < lim) i++;
          fine.push_back((uint16_t)(i - first[r]));
        }
      }
      if (fine.size() >= (1u << 27)) fits = false;
      if (fits && !rank_setup(op, S, keys)) return false;
      if (fits) {
        if (hipMalloc((void **)&S->d_region, region.size() * sizeof(uint2)) != hipSuccess ||
            hipMalloc((void **)&S->d_fine, fine.size() * sizeof(uint16_t)) != hipSuccess)
          return false;
        if (hipMemcpyAsync(S->d_region, region.data(), region.size() * sizeof(uint2), hipMemcpyHostToDevice,
                           ctx->stream) != hipSuccess ||
            hipMemcpyAsync(S->d_fine, fine.data(), fine.size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                           ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
          return false;
        A.rshift = RS;
        A.region = S->d_region;
        A.fine = S->d_fine;
      }
    }
  }
  // balanced schedule: chunks of whole rows with at most BAL_CAP targets
  A.chunk = nullptr;
  A.nchunk = 0;
  // runs of 512 chunks per XCD (bal_chunk): against plain round-robin order
  // 1.352 -> 1.294 ms at L = 32 (runs of 32: 1.307, 256: 1.297, 1024: 1.296)
  A.xcdg = 512;
  if (A.region) {
    // 16-bit counts: at most 3 targets per site, family and row (L <= 32)
    uint16_t *d_cnt = nullptr;
    std::vector<uint16_t> cnt(b->D);
    if (hipMalloc((void **)&d_cnt, b->D * sizeof(uint16_t)) != hipSuccess) return false;
    k_row_targets<uint16_t><<<(unsigned)((b->D + 255) / 256), 256, 0, ctx->stream>>>(A, d_cnt);
    bool okc = hipMemcpyAsync(cnt.data(), d_cnt, b->D * sizeof(uint16_t), hipMemcpyDeviceToHost, ctx->stream) ==
                   hipSuccess &&
               hipStreamSynchronize(ctx->stream) == hipSuccess;
    (void)hipFree(d_cnt);
    if (!okc) return false;
    std::vector<uint32_t> cs;
    bool fit = true;
    uint32_t acc = 0, rows = 0;
    cs.push_back(0);
    for (int64_t i = 0; i < b->D; i++) {
      if (cnt[i] > BAL_CAP) fit = false;
      if (rows == BAL_TPB || acc + cnt[i] > BAL_CAP) {
        cs.push_back((uint32_t)i);
        acc = 0;
        rows = 0;
      }
      acc += cnt[i];
      rows++;
    }
    cs.push_back((uint32_t)b->D);
    if (fit) {
      if (hipMalloc((void **)&S->d_chunk, cs.size() * sizeof(uint32_t)) != hipSuccess) return false;
      if (hipMemcpyAsync(S->d_chunk, cs.data(), cs.size() * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream) !=
              hipSuccess ||
          hipStreamSynchronize(ctx->stream) != hipSuccess)
        return false;
      A.chunk = S->d_chunk;
      A.nchunk = (uint32_t)cs.size() - 1;
    }
  }
  return true;
}

template <int KIND, int PH, int NB, int LT>
static void launch_fast_nb(const FastArgs &A, int dtype, void *y, const void *x, int acc, hipStream_t st) {
  unsigned grid = (unsigned)((A.D + FAST_TPB - 1) / FAST_TPB);
  if (dtype == EDK_C64)
    k_matvec_fast<KIND, PH, NB, LT, double2, double>
        <<<grid, FAST_TPB, 0, st>>>(A, (const double2 *)x, (double2 *)y, acc);
  else
    k_matvec_fast<KIND, PH, NB, LT, float2, float><<<grid, FAST_TPB, 0, st>>>(A, (const float2 *)x, (float2 *)y, acc);
}

template <int KIND, int PH, int LT>
static void launch_fast_lt(const FastArgs &A, int dtype, void *y, const void *x, int acc, hipStream_t st) {
  if (A.batch >= 8) launch_fast_nb<KIND, PH, 8, LT>(A, dtype, y, x, acc, st);
  else if (A.batch >= 4) launch_fast_nb<KIND, PH, 4, LT>(A, dtype, y, x, acc, st);
  else launch_fast_nb<KIND, PH, 2, LT>(A, dtype, y, x, acc, st);
}

template <int KIND, int PH, int LT>
static void launch_bal_lt(const FastArgs &A, int dtype, void *y, const void *x, int acc, hipStream_t st) {
  if (dtype == EDK_C64)
    k_matvec_bal<KIND, PH, LT, double2, double><<<A.nchunk, BAL_TPB, 0, st>>>(A, (const double2 *)x, (double2 *)y, acc);
  else
    k_matvec_bal<KIND, PH, LT, float2, float><<<A.nchunk, BAL_TPB, 0, st>>>(A, (const float2 *)x, (float2 *)y, acc);
}

template <int KIND, int PH>
static void launch_fast(const FastArgs &A, int dtype, void *y, const void *x, int acc, hipStream_t st) {
  if (A.chunk) {
    // the rank directory when it applies (real phases, run-length canonicalisers)
    if (KIND != EDK_PROJECTED && PH != 2 && edk_fast_launch_rk(A, dtype, y, x, acc, st)) return;
    // rotation flavour: L = 32 (v_alignbit), 16 <= L < 32 with unit cell 1
    // (run-length canonicaliser on L-bit words), anything else (full search)
    if (A.L == 32 && A.A == 1) launch_bal_lt<KIND, PH, 32>(A, dtype, y, x, acc, st);
    else if (A.A == 1 && A.L >= 16 && KIND != EDK_PROJECTED) launch_bal_lt<KIND, PH, 1>(A, dtype, y, x, acc, st);
    else launch_bal_lt<KIND, PH, 0>(A, dtype, y, x, acc, st);
    return;
  }
  if (A.L == 32 && A.A == 1) launch_fast_lt<KIND, PH, 32>(A, dtype, y, x, acc, st);
  else launch_fast_lt<KIND, PH, 0>(A, dtype, y, x, acc, st);
}

bool edk_matvec_prepare(edk_op *op) {
  const bool f = !op->disable_fast && fast_setup(op);
  if (!f) {
    if (!op->disable_fast) (void)edk_perm_prepare(op);
    (void)edk_generic_setup(op);
  }
  return f;
}

// Launch one schedule: EDK_VARIANT_TILE is the balanced chunk kernel (one
// block per chunk of whole rows), EDK_VARIANT_LANE{2,4,8} the thread-per-row
// kernel with 2/4/8 targets in flight per lane.  Blocks of vectors have the
// tile form only.  false: the variant does not apply (the caller falls back
// to the generic kernel).
static bool launch_variant(edk_op *op, int dtype, void *y, const void *x, int64_t nb, int accumulate,
                           hipStream_t st, int var) {
  if (var == EDK_VARIANT_GENERIC) return false;
  FastState *S = (FastState *)op->fast;
  FastArgs A = S->args;
  if (nb != 1) {
    if (nb > 8 || var > EDK_VARIANT_TILE) return false;
    if (!edk_fast_launch_blk(A, dtype, y, x, (int)nb, accumulate, st)) return false;
    return hipGetLastError() == hipSuccess;
  }
  if (var == EDK_VARIANT_TILE && !A.chunk) return false;
  if (var >= EDK_VARIANT_LANE2) {
    A.chunk = nullptr;
    A.batch = var == EDK_VARIANT_LANE2 ? 2 : (var == EDK_VARIANT_LANE4 ? 4 : 8);
  }
  // phase mode: 0 all coefficients real and positive, 1 real signs, 2 complex
  switch (A.kind) {
    case EDK_TRANSLATION_PARITY:
      if (A.kp == 1 && A.P == 1) launch_fast<EDK_TRANSLATION_PARITY, 0>(A, dtype, y, x, accumulate, st);
      else launch_fast<EDK_TRANSLATION_PARITY, 1>(A, dtype, y, x, accumulate, st);
      break;
    case EDK_TRANSLATIONAL:
      if (A.kp == 0) launch_fast<EDK_TRANSLATIONAL, 0>(A, dtype, y, x, accumulate, st);
      else if (2 * A.kp == A.len) launch_fast<EDK_TRANSLATIONAL, 1>(A, dtype, y, x, accumulate, st);
      else launch_fast<EDK_TRANSLATIONAL, 2>(A, dtype, y, x, accumulate, st);
      break;
    case EDK_PROJECTED: launch_fast<EDK_PROJECTED, 0>(A, dtype, y, x, accumulate, st); break;
    default: return false;
  }
  return hipGetLastError() == hipSuccess;
}

bool edk_matvec_fast(edk_op *op, int dtype, void *y, const void *x, int64_t nb, int accumulate, hipStream_t st) {
  if (op->disable_fast) return false;
  if (!fast_setup(op)) return false;
  FastState *S = (FastState *)op->fast;
  const int var = (nb >= 1 && nb <= 8) ? S->tuned[dtype == EDK_C64 ? 0 : 1][nb] : EDK_VARIANT_AUTO;
  return launch_variant(op, dtype, y, x, nb, accumulate, st, var);
}

int edk_fast_set_variant(edk_op *op, int dtype, int64_t nb, int variant) {
  if (nb < 1 || nb > 8 || variant < EDK_VARIANT_AUTO || variant > EDK_VARIANT_LANE8)
    return edk_fail(EDK_ERR_ARG, "variant or block width out of range");
  if (variant >= EDK_VARIANT_LANE2 && nb != 1)
    return edk_fail(EDK_ERR_ARG, "lane schedules apply to one vector only");
  FastState *S = fast_state(op);
  // the Abelian two-site kernel has the tile form for one vector only
  const bool perm = variant == EDK_VARIANT_TILE && nb == 1 && !op->disable_fast && edk_perm_prepare(op);
  if (variant != EDK_VARIANT_AUTO && variant != EDK_VARIANT_GENERIC && !edk_matvec_prepare(op) && !perm)
    return edk_fail(EDK_ERR_UNSUPPORTED, "operator/basis not eligible for the specialised kernels");
  S->tuned[dtype == EDK_C64 ? 0 : 1][nb] = variant;
  return EDK_OK;
}

// Time every applicable schedule on scratch vectors (reps launches each, HIP
// events on the context stream) and keep the fastest for (dtype, nb).
int edk_fast_autotune(edk_op *op, int dtype, int64_t nb, int reps, int *variant, double *ms_out) {
  if (nb < 1 || nb > 8) return edk_fail(EDK_ERR_ARG, "autotune: block width must be 1..8");
  if (reps < 1) reps = 1;
  edk_ctx *ctx = op->ctx;
  const int64_t D = op->basis->D;
  const size_t bytes = (size_t)std::max<int64_t>(D, 1) * nb * (dtype == EDK_C64 ? 16 : 8);
  std::vector<int> cand = {EDK_VARIANT_GENERIC};
  if (edk_matvec_prepare(op)) {
    FastState *S = (FastState *)op->fast;
    if (S->args.chunk) cand.push_back(EDK_VARIANT_TILE);
    if (nb == 1) {
      cand.push_back(EDK_VARIANT_LANE2);
      cand.push_back(EDK_VARIANT_LANE4);
      cand.push_back(EDK_VARIANT_LANE8);
    }
  } else if (nb == 1 && !op->disable_fast && edk_perm_prepare(op)) {
    cand.push_back(EDK_VARIANT_TILE);
  }
  void *x = nullptr, *y = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipMalloc(&x, bytes) != hipSuccess || hipMalloc(&y, bytes) != hipSuccess) {
    (void)hipFree(x);
    return edk_fail(EDK_ERR_OOM, "autotune scratch vectors");
  }
  int rc = EDK_OK;
  if (hipMemsetAsync(x, 0, bytes, ctx->stream) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    rc = edk_fail(EDK_ERR_HIP, "autotune setup");
  FastState *S = fast_state(op);
  const int di = dtype == EDK_C64 ? 0 : 1;
  int best = EDK_VARIANT_GENERIC;
  double best_ms = 1e300;
  for (size_t c = 0; !rc && c < cand.size(); c++) {
    S->tuned[di][nb] = cand[c];
    if ((rc = edk_matvec_launch(op, dtype, y, x, nb, 0))) break;
    if (hipEventRecord(e0, ctx->stream) != hipSuccess) rc = edk_fail(EDK_ERR_HIP, "autotune event");
    for (int r = 0; !rc && r < reps; r++) rc = edk_matvec_launch(op, dtype, y, x, nb, 0);
    float ms = 0;
    if (!rc && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = edk_fail(EDK_ERR_HIP, "autotune timing");
    if (!rc && ms / reps < best_ms) {
      best_ms = ms / reps;
      best = cand[c];
    }
  }
  S->tuned[di][nb] = rc ? EDK_VARIANT_AUTO : best;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(x);
  (void)hipFree(y);
  if (rc) return rc;
  if (variant) *variant = best;
  if (ms_out) *ms_out = best_ms;
  return EDK_OK;
}

// ---------------------------------------------------------------------------
// Multi-coefficient launches (the server's batched matvecs)
// ---------------------------------------------------------------------------
bool edk_fast_signature(edk_op *op, uint64_t *sig) {
  if (op->disable_fast || !fast_setup(op)) return false;
  FastState *S = (FastState *)op->fast;
  if (!S->args.chunk || !S->d_coltab) return false;
  *sig = S->signature;
  return true;
}

int edk_fast_matvec_mc(edk_op *const *ops, int n, int dtype, const void *const *xs, void *const *ys, void *Xblk,
                       hipStream_t st) {
  if (n < 1 || n > MC_MAXV) return edk_fail(EDK_ERR_ARG, "edk_fast_matvec_mc: 1..8 operators");
  uint64_t sig0 = 0;
  if (!edk_fast_signature(ops[0], &sig0)) return edk_fail(EDK_ERR_UNSUPPORTED, "edk_fast_matvec_mc: no fast path");
  const int nv = n == 1 ? 1 : (n == 2 ? 2 : (n <= 4 ? 4 : 8));
  McArgs M{};
  ColPtrs G{};
  for (int c = 0; c < nv; c++) {
    edk_op *op = ops[c < n ? c : 0];
    uint64_t sg = 0;
    if (op->basis != ops[0]->basis || !edk_fast_signature(op, &sg) || sg != sig0)
      return edk_fail(EDK_ERR_ARG, "edk_fast_matvec_mc: operators differ in basis or structure");
    M.ctab[c] = ((FastState *)op->fast)->d_coltab;
    M.y[c] = c < n ? ys[c] : nullptr;
    G.p[c] = xs[c < n ? c : 0];
  }
  const FastArgs &A = ((FastState *)ops[0]->fast)->args;
  // one column reads its vector in place; several are gathered into the block
  const void *X = xs[0];
  if (nv > 1) {
    if (int rc = edk_gather_cols(G, nv, (int64_t)A.D, dtype, Xblk, st)) return rc;
    X = Xblk;
  }
  if (!edk_fast_launch_mc(A, M, dtype, X, nv, st)) return edk_fail(EDK_ERR_UNSUPPORTED, "edk_fast_matvec_mc");
  EDK_HIP(hipGetLastError());
  return EDK_OK;
}

