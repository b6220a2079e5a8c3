// edk_assembly.hip -- explicit sector matrix on the device: sparse(opt),
// Array(opt) and the sparse!(opt) cache (Operator.jl:171-285).
//
// The matrix is assembled row by row with the same enumeration and
// canonicalisation as the generic gather kernel: row i holds H[i, i] (the
// diagonal local entries) and, for every off-diagonal local entry whose
// target lies in the sector, H[i, j] += M[row, loc] * conj(phase) * R_j / R_i
// -- the element the gather kernel applies to x_j.  Each row's entries are
// sorted by column, duplicates are summed and exact zeros dropped; rows keep
// their slot ranges from the counting pass (offsets rowptr[i], merged length
// rowlen[i]), and the download compacts them.
//
// As in the reference, the cached matrix serves products with several
// right-hand sides (edk_matvec with nb > 1: a CSR SpMM kernel), while a
// single vector keeps the matrix-free kernels.
#include <algorithm>
#include <vector>

#include "edk_matvec.h"

struct edk_csr {
  int64_t D = 0, slots = 0, nnz = 0;
  uint64_t *rowptr = nullptr;  // D + 1 slot offsets
  uint32_t *rowlen = nullptr;  // merged entries per row
  uint32_t *col = nullptr;     // 0-based columns
  double2 *val = nullptr;
  float2 *valf = nullptr;      // the same values rounded to FP32 (ComplexF32 products)
};

// the target of local entry (t, c) of row state s: packed state and content delta
__device__ __forceinline__ uint64_t asm_target(const MatvecArgs &A, const int *td, int t, int c, uint64_t s,
                                               int64_t &delta) {
  const uint64_t fm = (1ull << A.w) - 1;
  const int ns = td[0];
  uint64_t s2 = s;
  delta = 0;
  int cc = c;
  for (int q = ns - 1; q >= 0; q--) {
    uint64_t nd = (uint64_t)(cc % A.base);
    cc /= A.base;
    int sh = td[2 + q];
    uint64_t od = (s2 >> sh) & fm;
    s2 = (s2 & ~(fm << sh)) | (nd << sh);
    delta += ((int64_t)nd - (int64_t)od) * (int64_t)A.T.tpw[t * EDK_MAXSITES + q];
  }
  return s2;
}

// Row i's entries, in generation order: (column, value).  MODE 0 counts,
// MODE 1 writes.  Returns the number of entries (diagonal slot included).
template <int MODE>
__device__ __forceinline__ uint32_t asm_row(const MatvecArgs &A, int64_t i, uint32_t *col, double2 *val) {
  const TermsDev &T = A.T;
  const uint64_t fm = (1ull << A.w) - 1;
  const uint64_t s = A.kind == EDK_TENSOR ? edk_content_to_packed_d((uint64_t)i, A.L, A.base, A.w) : A.S.keys[i];
  const double invR = (A.kind == EDK_TENSOR || A.kind == EDK_PROJECTED) ? 1.0 : 1.0 / A.R[i];
  double dr = 0.0, di = 0.0;
  uint32_t n = 1;  // slot 0: the diagonal
  for (int t = 0; t < T.nterms; t++) {
    const int *td = T.tdesc + t * EDK_TDESC;
    const int ns = td[0], m = td[1];
    int loc = 0;
    for (int q = 0; q < ns; q++) loc = loc * A.base + (int)((s >> td[2 + q]) & fm);
    const int *rp = T.rowptr + T.mrow_off[m];
    for (int e = rp[loc]; e < rp[loc + 1]; e++) {
      const int c = T.col[e];
      const double2 v = T.val[e];
      if (c == loc) {
        dr += v.x;
        di += v.y;
        continue;
      }
      int64_t delta;
      const uint64_t s2 = asm_target(A, td, t, c, s, delta);
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
        const double Rj = A.orbit_norms ? o.R : A.R[j];
        cr = o.phase.x * Rj;  // conj(phase) * R_j
        ci = -o.phase.y * Rj;
      }
      if (MODE == 1) {
        col[n] = (uint32_t)j;
        val[n] = make_double2((v.x * cr - v.y * ci) * invR, (v.x * ci + v.y * cr) * invR);
      }
      n++;
    }
  }
  if (MODE == 1) {
    col[0] = (uint32_t)i;
    val[0] = make_double2(dr, di);
  }
  return n;
}

__global__ void k_asm_count(MatvecArgs A, uint32_t *cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < A.D) cnt[i] = asm_row<0>(A, i, nullptr, nullptr);
}

// write, sort by column (insertion sort: rows hold tens of entries), merge
// duplicates in order, drop exact zeros
__global__ void k_asm_fill(MatvecArgs A, const uint64_t *rowptr, uint32_t *col, double2 *val, uint32_t *rowlen) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.D) return;
  uint32_t *c = col + rowptr[i];
  double2 *v = val + rowptr[i];
  const uint32_t n = asm_row<1>(A, i, c, v);
  for (uint32_t a = 1; a < n; a++) {
    const uint32_t ca = c[a];
    const double2 va = v[a];
    uint32_t b = a;
    while (b > 0 && c[b - 1] > ca) {
      c[b] = c[b - 1];
      v[b] = v[b - 1];
      b--;
    }
    c[b] = ca;
    v[b] = va;
  }
  uint32_t m = 0;
  for (uint32_t a = 0; a < n; a++) {
    if (m > 0 && c[m - 1] == c[a]) {
      v[m - 1].x += v[a].x;
      v[m - 1].y += v[a].y;
    } else {
      c[m] = c[a];
      v[m] = v[a];
      m++;
    }
  }
  uint32_t k = 0;
  for (uint32_t a = 0; a < m; a++) {
    if (v[a].x != 0.0 || v[a].y != 0.0) {
      c[k] = c[a];
      v[k] = v[a];
      k++;
    }
  }
  rowlen[i] = k;
}

__global__ void k_asm_round_f32(const double2 *val, float2 *valf, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) valf[k] = make_float2((float)val[k].x, (float)val[k].y);
}

// Y = (accumulate ? Y : 0) + H X with NV consecutive lanes per row (one per
// column), entries summed in column order.  HT is the stored value type:
// FP64 for ComplexF64 vectors, FP32 for ComplexF32 vectors.
template <typename VT, typename ST, typename HT>
__global__ void __launch_bounds__(256) k_csr_spmm(const uint64_t *rowptr, const uint32_t *rowlen, const uint32_t *col,
                                                  const HT *val, int64_t D, int nv, const VT *__restrict__ x,
                                                  VT *__restrict__ y, int accumulate) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= D * nv) return;
  const int64_t i = g / nv;
  const int c = (int)(g % nv);
  const uint64_t e0 = rowptr[i];
  const uint32_t n = rowlen[i];
  ST ar = 0, ai = 0;
  // eight entries' loads in flight, added in order
  uint32_t e = 0;
  for (; e + 8 <= n; e += 8) {
    HT h[8];
    VT xj[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      h[u] = val[e0 + e + u];
      xj[u] = x[(size_t)col[e0 + e + u] * nv + c];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      ar += (ST)h[u].x * (ST)xj[u].x - (ST)h[u].y * (ST)xj[u].y;
      ai += (ST)h[u].x * (ST)xj[u].y + (ST)h[u].y * (ST)xj[u].x;
    }
  }
  for (; e < n; e++) {
    const uint32_t j = col[e0 + e];
    const HT h = val[e0 + e];
    const VT xj = x[(size_t)j * nv + c];
    ar += (ST)h.x * (ST)xj.x - (ST)h.y * (ST)xj.y;
    ai += (ST)h.x * (ST)xj.y + (ST)h.y * (ST)xj.x;
  }
  VT out;
  if (accumulate) {
    const VT y0 = y[g];
    out.x = y0.x + ar;
    out.y = y0.y + ai;
  } else {
    out.x = ar;
    out.y = ai;
  }
  y[g] = out;
}

void edk_csr_release(edk_op *op) {
  edk_csr *M = (edk_csr *)op->csr;
  if (!M) return;
  (void)hipFree(M->rowptr);
  (void)hipFree(M->rowlen);
  (void)hipFree(M->col);
  (void)hipFree(M->val);
  (void)hipFree(M->valf);
  delete M;
  op->csr = nullptr;
}

bool edk_matvec_csr(edk_op *op, int dtype, void *y, const void *x, int64_t nb, int accumulate, hipStream_t st) {
  edk_csr *M = (edk_csr *)op->csr;
  if (!M || nb < 2 || M->D == 0) return false;
  const int64_t n = M->D * nb;
  const unsigned grid = (unsigned)((n + 255) / 256);
  if (dtype == EDK_C64)
    k_csr_spmm<double2, double, double2><<<grid, 256, 0, st>>>(M->rowptr, M->rowlen, M->col, M->val, M->D, (int)nb,
                                                              (const double2 *)x, (double2 *)y, accumulate);
  else
    k_csr_spmm<float2, float, float2><<<grid, 256, 0, st>>>(M->rowptr, M->rowlen, M->col, M->valf, M->D, (int)nb,
                                                           (const float2 *)x, (float2 *)y, accumulate);
  return hipGetLastError() == hipSuccess;
}

extern "C" int edk_operator_assemble(edk_op *op, int64_t *nnz) {
  if (!op) return edk_fail(EDK_ERR_ARG, "null operator");
  edk_ctx *ctx = op->ctx;
  EDK_HIP(hipSetDevice(ctx->device));
  if (op->csr) {
    if (nnz) *nnz = ((edk_csr *)op->csr)->nnz;
    return EDK_OK;
  }
  MatvecArgs a = op->args();
  if (a.D >= (1ll << 32) - 1) return edk_fail(EDK_ERR_UNSUPPORTED, "sector too large for 32-bit columns");
  edk_csr *M = new edk_csr();
  M->D = a.D;
  std::vector<uint32_t> cnt(a.D);
  std::vector<uint64_t> rp(a.D + 1, 0);
  uint32_t *d_cnt = nullptr;
  int rc = EDK_OK;
  const unsigned grid = (unsigned)((a.D + 255) / 256);
  if (a.D > 0) {
    if (hipMalloc((void **)&d_cnt, a.D * sizeof(uint32_t)) != hipSuccess) rc = edk_fail(EDK_ERR_OOM, "row counts");
    if (!rc) {
      k_asm_count<<<grid, 256, 0, ctx->stream>>>(a, d_cnt);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(cnt.data(), d_cnt, a.D * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
          hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = edk_fail(EDK_ERR_HIP, "assembly count pass");
    }
  }
  for (int64_t i = 0; !rc && i < a.D; i++) rp[i + 1] = rp[i] + cnt[i];
  M->slots = (int64_t)rp[a.D];
  if (!rc && (hipMalloc((void **)&M->rowptr, (a.D + 1) * sizeof(uint64_t)) != hipSuccess ||
              hipMalloc((void **)&M->rowlen, std::max<int64_t>(a.D, 1) * sizeof(uint32_t)) != hipSuccess ||
              hipMalloc((void **)&M->col, std::max<int64_t>(M->slots, 1) * sizeof(uint32_t)) != hipSuccess ||
              hipMalloc((void **)&M->val, std::max<int64_t>(M->slots, 1) * sizeof(double2)) != hipSuccess ||
              hipMalloc((void **)&M->valf, std::max<int64_t>(M->slots, 1) * sizeof(float2)) != hipSuccess))
    rc = edk_fail(EDK_ERR_OOM, "sector matrix does not fit in device memory");
  if (!rc && hipMemcpyAsync(M->rowptr, rp.data(), (a.D + 1) * sizeof(uint64_t), hipMemcpyHostToDevice,
                            ctx->stream) != hipSuccess)
    rc = edk_fail(EDK_ERR_HIP, "row offsets");
  if (!rc && a.D > 0) {
    k_asm_fill<<<grid, 256, 0, ctx->stream>>>(a, M->rowptr, M->col, M->val, M->rowlen);
    if (M->slots > 0)
      k_asm_round_f32<<<(unsigned)((M->slots + 255) / 256), 256, 0, ctx->stream>>>(M->val, M->valf, M->slots);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(cnt.data(), M->rowlen, a.D * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = edk_fail(EDK_ERR_HIP, "assembly fill pass");
  }
  (void)hipFree(d_cnt);
  if (rc) {
    op->csr = M;
    edk_csr_release(op);
    return rc;
  }
  int64_t total = 0;
  for (int64_t i = 0; i < a.D; i++) total += cnt[i];
  M->nnz = total;
  op->csr = M;
  if (nnz) *nnz = total;
  return EDK_OK;
}

extern "C" int edk_operator_download_csr(edk_op *op, int64_t *rowptr_1based, int64_t *col_1based, double *val) {
  if (!op || !rowptr_1based || !col_1based || !val) return edk_fail(EDK_ERR_ARG, "edk_operator_download_csr: null argument");
  edk_csr *M = (edk_csr *)op->csr;
  if (!M) return edk_fail(EDK_ERR_STATE, "no assembled matrix: call edk_operator_assemble first");
  edk_ctx *ctx = op->ctx;
  EDK_HIP(hipSetDevice(ctx->device));
  std::vector<uint64_t> rp(M->D + 1);
  std::vector<uint32_t> len(M->D), col(M->slots);
  std::vector<double2> v(M->slots);
  EDK_HIP(hipMemcpyAsync(rp.data(), M->rowptr, (M->D + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  if (M->D > 0)
    EDK_HIP(hipMemcpyAsync(len.data(), M->rowlen, M->D * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  if (M->slots > 0) {
    EDK_HIP(hipMemcpyAsync(col.data(), M->col, M->slots * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    EDK_HIP(hipMemcpyAsync(v.data(), M->val, M->slots * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
  }
  EDK_HIP(hipStreamSynchronize(ctx->stream));
  int64_t k = 0;
  rowptr_1based[0] = 1;
  for (int64_t i = 0; i < M->D; i++) {
    for (uint32_t e = 0; e < len[i]; e++, k++) {
      col_1based[k] = (int64_t)col[rp[i] + e] + 1;
      val[2 * k] = v[rp[i] + e].x;
      val[2 * k + 1] = v[rp[i] + e].y;
    }
    rowptr_1based[i + 1] = k + 1;
  }
  return EDK_OK;
}

extern "C" int edk_operator_release_csr(edk_op *op) {
  if (!op) return edk_fail(EDK_ERR_ARG, "null operator");
  (void)hipSetDevice(op->ctx->device);
  (void)hipStreamSynchronize(op->ctx->stream);
  edk_csr_release(op);
  return EDK_OK;
}
