// edk_lindblad.hip -- the many-body Lindblad generator on dense density
// matrices and the adaptive Arnoldi propagator around it
// (LiouvillianMap / LindbladArnoldiCache / lindblad_timeevolve,
// LindbladEvolution.jl:62-611).
//
// Generator.  For rho (d x d, column-major, i.e. vec(rho) of length d^2)
//   L(rho) = -i[H, rho] + sum_k L_k rho L_k^H - 1/2 {D, rho},  D = sum_k L_k^H L_k
//          = K1 rho + rho K2 + sum_k L_k rho L_k^H,  K1 = -iH - D/2,  K2 = iH - D/2
// (_liouvillian_apply!, LindbladEvolution.jl:114-136, applies the same terms
// one product at a time).  On the device it is three rocBLAS calls, all on
// the FP64 matrix cores:
//   R  = [rho; rho L_1^H; ...; rho L_n^H]   rho copied in, the lower blocks by
//                                           one strided-batched zgemm
//   Y  = rho K2                             zgemm
//   Y += [K1 | L_1 | ... | L_n] R           one zgemm of inner size (n+1) d
// so the 2 + 2n products of order d run as two large GEMMs and one batch.
//
// Propagator.  The controller of LindbladArnoldiCache (defect monitor,
// interval sampling with bisection, basis extension, restart from a new
// anchor) runs on the host over the small (m+1) x m Hessenberg matrix; the
// Arnoldi basis (d^2 x (m_max+1)) and every product with it stay in HBM.
// The reference keeps the Schur form Hm = Z T Z^H and evaluates
// exp(tau T) Z^H e1; here the same vector Z exp(tau T) Z^H e1 = exp(tau Hm) e1
// comes from a scaling-and-squaring Pade exponential of Hm itself, and the
// defect |h_{m+1,m}| |e_m^T exp(tau Hm) e1| ||rho_anchor|| is read off it.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <map>
#include <unordered_map>
#include <vector>

#include "edk_krylov.h"

using cplx = std::complex<double>;

#define LV_TILE 32

#define EDK_RBCHK(call)                                                                    \
  do {                                                                                     \
    rocblas_status st_ = (call);                                                           \
    if (st_ != rocblas_status_success)                                                     \
      return edk_fail(EDK_ERR_HIP, std::string("rocBLAS/rocSOLVER: ") + rocblas_status_to_string(st_)); \
  } while (0)

// ---------------------------------------------------------------------------
// device kernels
// ---------------------------------------------------------------------------
// K1 = -i H - D/2 and K2 = i H - D/2; K2 may alias D (each element is read
// before it is written, by the same lane)
__global__ void k_lv_make(double2 *K1, double2 *K2, const double2 *H, const double2 *D, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 h = H[i], dd = D[i];
  K1[i] = make_double2(h.y - 0.5 * dd.x, -h.x - 0.5 * dd.y);
  K2[i] = make_double2(-h.y - 0.5 * dd.x, h.x - 0.5 * dd.y);
}

// trace of a d x d matrix: one block, fixed summation order
__global__ void __launch_bounds__(256) k_lv_trace(const double2 *A, int64_t d, double2 *out) {
  __shared__ double2 sm[256];
  double2 acc = make_double2(0, 0);
  for (int64_t i = threadIdx.x; i < d; i += 256) {
    const double2 a = A[i * (d + 1)];
    acc.x += a.x;
    acc.y += a.y;
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sm[threadIdx.x].x += sm[threadIdx.x + s].x;
      sm[threadIdx.x].y += sm[threadIdx.x + s].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sm[0];
}

// Per 32 x 32 tile (bx, by): sum over the tile of |A[r,c] - conj(A[c,r])|^2
// into part[by * gx + bx]; with Hp != null also Hp[r,c] = (A[r,c] + conj(A[c,r]))/2.
// The transposed tile is staged through LDS so both reads are coalesced.
__global__ void __launch_bounds__(256) k_lv_herm(const double2 *A, int64_t d, double *part, double2 *Hp) {
  __shared__ double2 t[LV_TILE][LV_TILE + 1];
  __shared__ double sm[256];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 lanes
  const int64_t r0 = (int64_t)blockIdx.x * LV_TILE, c0 = (int64_t)blockIdx.y * LV_TILE;
  // stage tile (rows c0.., cols r0..) transposed: t[cc][rr] = A[c0+cc, r0+rr]
  for (int q = ty; q < LV_TILE; q += 8) {
    const int64_t row = c0 + tx, col = r0 + q;
    t[tx][q] = (row < d && col < d) ? A[col * d + row] : make_double2(0, 0);
  }
  __syncthreads();
  double acc = 0.0;
  for (int q = ty; q < LV_TILE; q += 8) {
    const int64_t row = r0 + tx, col = c0 + q;
    if (row < d && col < d) {
      const double2 a = A[col * d + row];
      const double2 b = t[q][tx];  // A[col, row]
      const double dr = a.x - b.x, di = a.y + b.y;
      acc += dr * dr + di * di;
      if (Hp) Hp[col * d + row] = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
    }
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = sm[0];
}

// sum of n partials in a fixed order (one block)
__global__ void __launch_bounds__(256) k_lv_sum(const double *part, int64_t n, double *out) {
  __shared__ double sm[256];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) acc += part[i];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sm[0];
}

__global__ void k_lv_cscale(double2 *y, double2 a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 v = y[i];
  y[i] = make_double2(a.x * v.x - a.y * v.y, a.x * v.y + a.y * v.x);
}

// Sparse products with a dense column-major d x d X (element (r, c) at r + c d).
// One lane per (r, c), r fastest; every lane sums its terms in a fixed order.
//   left:  Y[r, c] (+)= sum_e v[e] X[ci[e], c] over row r of A   (Y = A X)
//   right: Y[r, c] += sum_e X[r, ci[e]] f(v[e]) over row c of M  (Y += X M^T, f = conj: X M^H)
// The right form reads row c of M uniformly across the block and X[r, k]
// coalesced; the left form gathers within the contiguous column c.
__global__ void __launch_bounds__(256) k_lv_spmm_left(double2 *__restrict__ Y, const int *__restrict__ rp,
                                                      const int *__restrict__ ci, const double2 *__restrict__ v,
                                                      const double2 *__restrict__ X, int64_t d, int accumulate) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= d) return;
  const int64_t c = blockIdx.y;
  const double2 *xc = X + c * d;
  double ar = 0.0, ai = 0.0;
  // four terms' loads in flight, added in order
  const int e1 = rp[r + 1];
  int e = rp[r];
  for (; e + 4 <= e1; e += 4) {
    double2 a[4], x[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      a[u] = v[e + u];
      x[u] = xc[ci[e + u]];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      ar += a[u].x * x[u].x - a[u].y * x[u].y;
      ai += a[u].x * x[u].y + a[u].y * x[u].x;
    }
  }
  for (; e < e1; e++) {
    const double2 a = v[e], x = xc[ci[e]];
    ar += a.x * x.x - a.y * x.y;
    ai += a.x * x.y + a.y * x.x;
  }
  double2 *yp = Y + r + c * d;
  if (accumulate) {
    const double2 y0 = *yp;
    ar += y0.x;
    ai += y0.y;
  }
  *yp = make_double2(ar, ai);
}

__global__ void __launch_bounds__(256) k_lv_spmm_right(double2 *__restrict__ Y, const int *__restrict__ rp,
                                                       const int *__restrict__ ci, const double2 *__restrict__ v,
                                                       const double2 *__restrict__ X, int64_t d, int conj) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= d) return;
  const int64_t c = blockIdx.y;
  const double sg = conj ? -1.0 : 1.0;
  double ar = 0.0, ai = 0.0;
  // four terms' loads in flight, added in order
  const int e1 = rp[c + 1];
  int e = rp[c];
  for (; e + 4 <= e1; e += 4) {
    double2 m[4], x[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      m[u] = v[e + u];
      x[u] = X[r + (int64_t)ci[e + u] * d];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const double mi = sg * m[u].y;
      ar += x[u].x * m[u].x - x[u].y * mi;
      ai += x[u].x * mi + x[u].y * m[u].x;
    }
  }
  for (; e < e1; e++) {
    const double2 m = v[e], x = X[r + (int64_t)ci[e] * d];
    const double mi = sg * m.y;
    ar += x.x * m.x - x.y * mi;
    ai += x.x * mi + x.y * m.x;
  }
  double2 *yp = Y + r + c * d;
  const double2 y0 = *yp;
  *yp = make_double2(y0.x + ar, y0.y + ai);
}

// ---------------------------------------------------------------------------
// LiouvillianMap
// ---------------------------------------------------------------------------
struct edk_liouvillian {
  edk_ctx *ctx = nullptr;
  int64_t d = 0;
  int nj = 0;
  // dense form
  double2 *Lb = nullptr;  // [K1 | L_1 | ... | L_nj]: d x (nj+1) d, column-major
  double2 *K2 = nullptr;  // i H - D/2
  double2 *Rb = nullptr;  // [X; X L_1^H; ...; X L_nj^H]: (nj+1) d x d
  // sparse form: CSR of K1, of K2^T and of every jump (rows of all
  // matrices back to back: matrix q has rows rp[q (d+1) ...])
  bool sparse = false;
  int *rp = nullptr, *ci = nullptr;
  double2 *cv = nullptr;
  double2 *T = nullptr;   // L_k X
  rocblas_handle h = nullptr;
};

static void lv_free(edk_liouvillian *A) {
  if (!A) return;
  if (A->h) (void)rocblas_destroy_handle(A->h);
  edk_dfree(A->ctx, A->Lb);
  edk_dfree(A->ctx, A->K2);
  edk_dfree(A->ctx, A->Rb);
  edk_dfree(A->ctx, A->rp);
  edk_dfree(A->ctx, A->ci);
  edk_dfree(A->ctx, A->cv);
  edk_dfree(A->ctx, A->T);
  delete A;
}

static const rocblas_double_complex *rbc(const void *p) { return (const rocblas_double_complex *)p; }
static rocblas_double_complex *rbm(void *p) { return (rocblas_double_complex *)p; }

// y = L(x): x, y device d x d (column-major), distinct
// sparse form: Y = K1 X + X K2 + sum_k (L_k X) L_k^H, 2 + 2 nj passes over d^2
static int lv_apply_sparse(edk_liouvillian *A, const void *x, void *y) {
  const int64_t d = A->d, d1 = d + 1;
  hipStream_t s = A->ctx->stream;
  const dim3 grid((unsigned)((d + 255) / 256), (unsigned)d);
  const double2 *X = (const double2 *)x;
  double2 *Y = (double2 *)y;
  k_lv_spmm_left<<<grid, 256, 0, s>>>(Y, A->rp, A->ci, A->cv, X, d, 0);
  k_lv_spmm_right<<<grid, 256, 0, s>>>(Y, A->rp + d1, A->ci, A->cv, X, d, 0);
  for (int k = 0; k < A->nj; k++) {
    const int *rk = A->rp + (2 + k) * d1;
    k_lv_spmm_left<<<grid, 256, 0, s>>>(A->T, rk, A->ci, A->cv, X, d, 0);
    k_lv_spmm_right<<<grid, 256, 0, s>>>(Y, rk, A->ci, A->cv, A->T, d, 1);
  }
  EDK_HIP(hipGetLastError());
  return EDK_OK;
}

static int lv_apply(edk_liouvillian *A, const void *x, void *y) {
  if (A->sparse) return lv_apply_sparse(A, x, y);
  const int64_t d = A->d, dd = d * d;
  const int nj = A->nj;
  const rocblas_int n = (rocblas_int)d, ldr = (rocblas_int)((nj + 1) * d);
  const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
  hipStream_t s = A->ctx->stream;
  const void *R = x;
  rocblas_int ldR = n;
  if (nj > 0) {
    EDK_HIP(hipMemcpy2DAsync(A->Rb, (size_t)ldr * sizeof(double2), x, (size_t)d * sizeof(double2),
                             (size_t)d * sizeof(double2), (size_t)d, hipMemcpyDeviceToDevice, s));
    // R block k (rows k d .. (k+1) d) = X L_k^H, k = 1..nj
    EDK_RBCHK(rocblas_zgemm_strided_batched(A->h, rocblas_operation_none, rocblas_operation_conjugate_transpose, n, n,
                                            n, &one, rbc(x), n, 0, rbc(A->Lb + dd), n, dd, &zero, rbm(A->Rb + d), ldr,
                                            d, nj));
    R = A->Rb;
    ldR = ldr;
  }
  EDK_RBCHK(rocblas_zgemm(A->h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, rbc(x), n, rbc(A->K2),
                          n, &zero, rbm(y), n));
  EDK_RBCHK(rocblas_zgemm(A->h, rocblas_operation_none, rocblas_operation_none, n, n, (rocblas_int)((nj + 1) * d), &one,
                          rbc(A->Lb), n, rbc(R), ldR, &one, rbm(y), n));
  return EDK_OK;
}

// LiouvillianMap(H, jumps) (LindbladEvolution.jl:78-92); H and the nj jumps
// are device d x d complex128 column-major matrices (jumps one after another)
extern "C" int edk_liouvillian_create(edk_ctx *ctx, int64_t d, const void *H, int nj, const void *jumps,
                                      edk_liouvillian **out) {
  if (!ctx || !H || !out || (nj > 0 && !jumps)) return edk_fail(EDK_ERR_ARG, "edk_liouvillian_create: null argument");
  if (d < 1) return edk_fail(EDK_ERR_ARG, "LiouvillianMap: H must be non-empty");
  if (nj < 0) return edk_fail(EDK_ERR_ARG, "LiouvillianMap: negative number of jumps");
  if (d > 46340) return edk_fail(EDK_ERR_UNSUPPORTED, "LiouvillianMap: dimension above 46340");
  if ((int64_t)(nj + 1) * d > (int64_t)INT32_MAX) return edk_fail(EDK_ERR_UNSUPPORTED, "LiouvillianMap: too many jumps");
  EDK_HIP(hipSetDevice(ctx->device));
  const int64_t dd = d * d;
  const size_t mb = (size_t)dd * sizeof(double2);
  edk_liouvillian *A = new edk_liouvillian();
  A->ctx = ctx;
  A->d = d;
  A->nj = nj;
  int rc = EDK_OK;
  if (edk_dalloc(ctx, (void **)&A->Lb, mb * (nj + 1)) != hipSuccess || edk_dalloc(ctx, (void **)&A->K2, mb) != hipSuccess ||
      (nj > 0 && edk_dalloc(ctx, (void **)&A->Rb, mb * (nj + 1)) != hipSuccess))
    rc = edk_fail(EDK_ERR_OOM, "LiouvillianMap does not fit in device memory");
  if (!rc && rocblas_create_handle(&A->h) != rocblas_status_success) {
    A->h = nullptr;
    rc = edk_fail(EDK_ERR_HIP, "rocblas_create_handle");
  }
  if (!rc && rocblas_set_stream(A->h, ctx->stream) != rocblas_status_success)
    rc = edk_fail(EDK_ERR_HIP, "rocblas_set_stream");
  hipStream_t s = ctx->stream;
  if (!rc && nj > 0 && hipMemcpyAsync(A->Lb + dd, jumps, mb * nj, hipMemcpyDeviceToDevice, s) != hipSuccess)
    rc = edk_fail(EDK_ERR_HIP, "copy of the jump operators");
  // D = sum_k L_k^H L_k, accumulated in K2's buffer
  if (!rc && nj == 0 && hipMemsetAsync(A->K2, 0, mb, s) != hipSuccess) rc = edk_fail(EDK_ERR_HIP, "clear D");
  const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
  for (int k = 0; k < nj && !rc; k++) {
    const double2 *Lk = A->Lb + dd * (k + 1);
    if (rocblas_zgemm(A->h, rocblas_operation_conjugate_transpose, rocblas_operation_none, (rocblas_int)d,
                      (rocblas_int)d, (rocblas_int)d, &one, rbc(Lk), (rocblas_int)d, rbc(Lk), (rocblas_int)d,
                      k == 0 ? &zero : &one, rbm(A->K2), (rocblas_int)d) != rocblas_status_success)
      rc = edk_fail(EDK_ERR_HIP, "rocBLAS zgemm (D = sum L^H L)");
  }
  if (!rc) {
    k_lv_make<<<(unsigned)((dd + 255) / 256), 256, 0, s>>>(A->Lb, A->K2, (const double2 *)H, A->K2, dd);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
      rc = edk_fail(EDK_ERR_HIP, "LiouvillianMap set-up");
  }
  if (rc) {
    lv_free(A);
    return rc;
  }
  *out = A;
  return EDK_OK;
}

// LiouvillianMap(sparse(H), sparse.(jumps)) (LindbladEvolution.jl:78-92 with
// issparse): host CSR arrays, 0-based, complex values interleaved.  H has
// rows h_rp[0..d]; jump q has rows j_rp[q (d+1) ...] with offsets into
// j_col / j_val.  D = sum_k L_k^H L_k, K1 = -iH - D/2 and K2^T are formed
// here (sparse), then everything is copied to the device.
extern "C" int edk_liouvillian_create_csr(edk_ctx *ctx, int64_t d, const int64_t *h_rp, const int32_t *h_col,
                                          const double *h_val, int nj, const int64_t *j_rp, const int32_t *j_col,
                                          const double *j_val, edk_liouvillian **out) {
  if (!ctx || !h_rp || !out || (nj > 0 && !j_rp)) return edk_fail(EDK_ERR_ARG, "edk_liouvillian_create_csr: null argument");
  if (d < 1) return edk_fail(EDK_ERR_ARG, "LiouvillianMap: H must be non-empty");
  if (nj < 0) return edk_fail(EDK_ERR_ARG, "LiouvillianMap: negative number of jumps");
  if (d > 46340) return edk_fail(EDK_ERR_UNSUPPORTED, "LiouvillianMap: dimension above 46340");
  using Row = std::map<int, cplx>;
  auto check_rows = [&](const int64_t *rp, const int32_t *col, const char *what) -> int {
    if (rp[0] < 0) return edk_fail(EDK_ERR_ARG, std::string(what) + ": negative row pointer");
    for (int64_t r = 0; r < d; r++) {
      if (rp[r + 1] < rp[r]) return edk_fail(EDK_ERR_ARG, std::string(what) + ": decreasing row pointers");
      for (int64_t e = rp[r]; e < rp[r + 1]; e++)
        if (col[e] < 0 || col[e] >= d) return edk_fail(EDK_ERR_ARG, std::string(what) + ": column out of range");
    }
    return EDK_OK;
  };
  if (int rc = check_rows(h_rp, h_col, "H")) return rc;
  for (int q = 0; q < nj; q++)
    if (int rc = check_rows(j_rp + (int64_t)q * (d + 1), j_col, "jump")) return rc;
  // D[i, j] = sum_k sum_r conj(L_k[r, i]) L_k[r, j]
  std::vector<Row> D(d), K1(d), K2T(d);
  for (int q = 0; q < nj; q++) {
    const int64_t *rp = j_rp + (int64_t)q * (d + 1);
    for (int64_t r = 0; r < d; r++)
      for (int64_t a = rp[r]; a < rp[r + 1]; a++) {
        const cplx la = std::conj(cplx(j_val[2 * a], j_val[2 * a + 1]));
        for (int64_t b = rp[r]; b < rp[r + 1]; b++) D[j_col[a]][j_col[b]] += la * cplx(j_val[2 * b], j_val[2 * b + 1]);
      }
  }
  for (int64_t r = 0; r < d; r++) {
    for (int64_t e = h_rp[r]; e < h_rp[r + 1]; e++) {
      const cplx h(h_val[2 * e], h_val[2 * e + 1]);
      K1[r][h_col[e]] += cplx(0, -1) * h;
      K2T[h_col[e]][r] += cplx(0, 1) * h;
    }
    for (auto &kv : D[r]) {
      K1[r][kv.first] += -0.5 * kv.second;
      K2T[kv.first][r] += -0.5 * kv.second;
    }
  }
  // flatten: K1, K2^T, then the jumps as given
  std::vector<int> rp;
  std::vector<int> ci;
  std::vector<double2> cv;
  auto push_rows = [&](const std::vector<Row> &M) {
    for (int64_t r = 0; r < d; r++) {
      rp.push_back((int)ci.size());
      for (auto &kv : M[r]) {
        ci.push_back(kv.first);
        cv.push_back(make_double2(kv.second.real(), kv.second.imag()));
      }
    }
    rp.push_back((int)ci.size());
  };
  size_t total = 0;
  for (int64_t r = 0; r < d; r++) total += K1[r].size() + K2T[r].size();
  for (int q = 0; q < nj; q++) total += (size_t)(j_rp[(int64_t)q * (d + 1) + d] - j_rp[(int64_t)q * (d + 1)]);
  if (total >= (size_t)INT32_MAX) return edk_fail(EDK_ERR_UNSUPPORTED, "LiouvillianMap: more than 2^31 stored entries");
  push_rows(K1);
  push_rows(K2T);
  for (int q = 0; q < nj; q++) {
    const int64_t *jr = j_rp + (int64_t)q * (d + 1);
    for (int64_t r = 0; r < d; r++) {
      rp.push_back((int)ci.size());
      for (int64_t e = jr[r]; e < jr[r + 1]; e++) {
        ci.push_back(j_col[e]);
        cv.push_back(make_double2(j_val[2 * e], j_val[2 * e + 1]));
      }
    }
    rp.push_back((int)ci.size());
  }
  EDK_HIP(hipSetDevice(ctx->device));
  edk_liouvillian *A = new edk_liouvillian();
  A->ctx = ctx;
  A->d = d;
  A->nj = nj;
  A->sparse = true;
  int rc = EDK_OK;
  const size_t nnz = std::max<size_t>(ci.size(), 1);
  if (edk_dalloc(ctx, (void **)&A->rp, rp.size() * sizeof(int)) != hipSuccess ||
      edk_dalloc(ctx, (void **)&A->ci, nnz * sizeof(int)) != hipSuccess ||
      edk_dalloc(ctx, (void **)&A->cv, nnz * sizeof(double2)) != hipSuccess ||
      (nj > 0 && edk_dalloc(ctx, (void **)&A->T, (size_t)d * d * sizeof(double2)) != hipSuccess))
    rc = edk_fail(EDK_ERR_OOM, "LiouvillianMap does not fit in device memory");
  if (!rc && rocblas_create_handle(&A->h) != rocblas_status_success) {
    A->h = nullptr;
    rc = edk_fail(EDK_ERR_HIP, "rocblas_create_handle");
  }
  if (!rc && rocblas_set_stream(A->h, ctx->stream) != rocblas_status_success)
    rc = edk_fail(EDK_ERR_HIP, "rocblas_set_stream");
  hipStream_t st = ctx->stream;
  if (!rc && (hipMemcpyAsync(A->rp, rp.data(), rp.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
              (!ci.empty() && hipMemcpyAsync(A->ci, ci.data(), ci.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess) ||
              (!cv.empty() && hipMemcpyAsync(A->cv, cv.data(), cv.size() * sizeof(double2), hipMemcpyHostToDevice, st) != hipSuccess) ||
              hipStreamSynchronize(st) != hipSuccess))
    rc = edk_fail(EDK_ERR_HIP, "copy of the sparse Liouvillian");
  if (rc) {
    lv_free(A);
    return rc;
  }
  *out = A;
  return EDK_OK;
}

// A * rho (LindbladEvolution.jl:138-151): y = L(x), device d x d buffers
extern "C" int edk_liouvillian_apply(edk_liouvillian *A, const void *x, void *y) {
  if (!A || !x || !y) return edk_fail(EDK_ERR_ARG, "edk_liouvillian_apply: null argument");
  if (x == y) return edk_fail(EDK_ERR_ARG, "edk_liouvillian_apply: input and output must differ");
  EDK_HIP(hipSetDevice(A->ctx->device));
  if (int rc = lv_apply(A, x, y)) return rc;
  EDK_HIP(hipStreamSynchronize(A->ctx->stream));
  return EDK_OK;
}

extern "C" int edk_liouvillian_destroy(edk_liouvillian *A) {
  if (!A) return EDK_OK;
  hipSetDevice(A->ctx->device);
  hipStreamSynchronize(A->ctx->stream);
  lv_free(A);
  return EDK_OK;
}

// ---------------------------------------------------------------------------
// small dense complex matrices on the host (the reduced model)
// ---------------------------------------------------------------------------
// C = A B, n x n column-major
static void cmm(int n, const cplx *A, const cplx *B, cplx *C) {
  std::fill(C, C + (size_t)n * n, cplx(0));
  for (int j = 0; j < n; j++)
    for (int k = 0; k < n; k++) {
      const cplx b = B[k + (size_t)j * n];
      if (b == cplx(0)) continue;
      const cplx *a = A + (size_t)k * n;
      cplx *c = C + (size_t)j * n;
      for (int i = 0; i < n; i++) c[i] += a[i] * b;
    }
}

// largest absolute column sum (opnorm(., 1))
static double onenorm(int n, const cplx *A) {
  double m = 0;
  for (int j = 0; j < n; j++) {
    double s = 0;
    for (int i = 0; i < n; i++) s += std::abs(A[i + (size_t)j * n]);
    m = std::max(m, s);
  }
  return m;
}

// Solve M X = B in place (B := X), Gaussian elimination with partial
// pivoting; a zero pivot fills X with NaN.
static void csolve(int n, std::vector<cplx> &M, std::vector<cplx> &B) {
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = std::abs(M[k + (size_t)k * n]);
    for (int i = k + 1; i < n; i++)
      if (std::abs(M[i + (size_t)k * n]) > best) best = std::abs(M[i + (size_t)k * n]), p = i;
    if (!(best > 0)) {
      std::fill(B.begin(), B.end(), cplx(NAN, NAN));
      return;
    }
    if (p != k) {
      for (int j = 0; j < n; j++) {
        std::swap(M[k + (size_t)j * n], M[p + (size_t)j * n]);
        std::swap(B[k + (size_t)j * n], B[p + (size_t)j * n]);
      }
    }
    const cplx piv = M[k + (size_t)k * n];
    for (int i = k + 1; i < n; i++) {
      const cplx f = M[i + (size_t)k * n] / piv;
      if (f == cplx(0)) continue;
      for (int j = k; j < n; j++) M[i + (size_t)j * n] -= f * M[k + (size_t)j * n];
      for (int j = 0; j < n; j++) B[i + (size_t)j * n] -= f * B[k + (size_t)j * n];
    }
  }
  for (int j = 0; j < n; j++)
    for (int i = n - 1; i >= 0; i--) {
      cplx s = B[i + (size_t)j * n];
      for (int k = i + 1; k < n; k++) s -= M[i + (size_t)k * n] * B[k + (size_t)j * n];
      B[i + (size_t)j * n] = s / M[i + (size_t)i * n];
    }
}

// E = exp(A), n x n: scaling and squaring with the diagonal [13/13] Pade
// approximant (N. J. Higham, "The scaling and squaring method for the matrix
// exponential revisited", SIAM J. Matrix Anal. Appl. 26 (2005) 1179-1193,
// Algorithm 2.3, degree 13 at every norm).
static void cexpm(int n, const cplx *A, std::vector<cplx> &E) {
  static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0, 1187353796428800.0,
                               129060195264000.0,   10559470521600.0,    670442572800.0,    33522128640.0,
                               1323241920.0,        40840800.0,          960960.0,          16380.0,
                               182.0,               1.0};
  const double theta13 = 5.371920351148152;
  const size_t nn = (size_t)n * n;
  E.assign(nn, cplx(0));
  const double nrm = onenorm(n, A);
  if (!std::isfinite(nrm)) {
    std::fill(E.begin(), E.end(), cplx(NAN, NAN));
    return;
  }
  int s = 0;
  if (nrm > theta13) s = (int)std::ceil(std::log2(nrm / theta13));
  const double sc = std::ldexp(1.0, -s);
  std::vector<cplx> X(nn), X2(nn), X4(nn), X6(nn), T(nn), U(nn), V(nn);
  for (size_t i = 0; i < nn; i++) X[i] = A[i] * sc;
  cmm(n, X.data(), X.data(), X2.data());
  cmm(n, X2.data(), X2.data(), X4.data());
  cmm(n, X4.data(), X2.data(), X6.data());
  // U = X [X6 (b13 X6 + b11 X4 + b9 X2) + b7 X6 + b5 X4 + b3 X2 + b1 I]
  for (size_t i = 0; i < nn; i++) T[i] = b[13] * X6[i] + b[11] * X4[i] + b[9] * X2[i];
  cmm(n, X6.data(), T.data(), U.data());
  for (size_t i = 0; i < nn; i++) U[i] += b[7] * X6[i] + b[5] * X4[i] + b[3] * X2[i];
  for (int i = 0; i < n; i++) U[i + (size_t)i * n] += b[1];
  cmm(n, X.data(), U.data(), T.data());
  U.swap(T);
  // V = X6 (b12 X6 + b10 X4 + b8 X2) + b6 X6 + b4 X4 + b2 X2 + b0 I
  for (size_t i = 0; i < nn; i++) T[i] = b[12] * X6[i] + b[10] * X4[i] + b[8] * X2[i];
  cmm(n, X6.data(), T.data(), V.data());
  for (size_t i = 0; i < nn; i++) V[i] += b[6] * X6[i] + b[4] * X4[i] + b[2] * X2[i];
  for (int i = 0; i < n; i++) V[i + (size_t)i * n] += b[0];
  // (V - U) R = (V + U)
  std::vector<cplx> M(nn), R(nn);
  for (size_t i = 0; i < nn; i++) {
    M[i] = V[i] - U[i];
    R[i] = V[i] + U[i];
  }
  csolve(n, M, R);
  for (int q = 0; q < s; q++) {
    cmm(n, R.data(), R.data(), T.data());
    R.swap(T);
  }
  E.swap(R);
}

// ---------------------------------------------------------------------------
// LindbladArnoldiCache
// ---------------------------------------------------------------------------
struct edk_lindblad {
  edk_liouvillian *A = nullptr;
  edk_ctx *ctx = nullptr;
  int64_t d = 0, N = 0;
  double t_anchor = 0, t_latest = 0;
  double2 *rho_anchor = nullptr;
  double norm_anchor = 0;
  cplx trace_target;
  double2 *V = nullptr;  // N x (m_max+1)
  double2 *w = nullptr;
  std::vector<cplx> Hred;  // (m_max+1) x m_max, column-major
  int m = 0;
  // reduced model: Hm = Hred[:m, :m], its 1-norm, exp(tau Hm) by tau
  std::vector<cplx> Hm;
  double gamma = 0;
  std::unordered_map<uint64_t, std::vector<cplx>> step_cache;
  edk_lindblad_params p;
  edk_lindblad_diag diag;
  std::vector<double> accepted, trace_drifts, herm_drifts, min_eigs;
  KrylovWork K;
  // diagnostics buffers
  double *part = nullptr;  // tile partials of the Hermiticity drift
  double *dsc = nullptr;   // [0]: drift^2, [1..2]: trace, then eigenvalues
  double2 *herm = nullptr; // Hermitian part for the smallest eigenvalue
  double *ework = nullptr;
  rocblas_int *info = nullptr;
  int64_t nparts = 0;
};

extern "C" void edk_lindblad_default_params(edk_lindblad_params *p) {
  p->tol = 1e-10;
  p->m_init = 20;
  p->m_max = 50;
  p->extend_step = 10;
  p->nsample = 17;
  p->bisect_iters = 12;
  p->extend_basis = 1;
  p->normalize_trace = 0;
  p->record_min_eig = 0;
}

static inline cplx &hred(edk_lindblad *c, int i, int j) { return c->Hred[i + (size_t)j * (c->p.m_max + 1)]; }

// _arnoldi_build! (LindbladEvolution.jl:282-321): Gram-Schmidt with the
// coefficients accumulated into the Hessenberg column (classical sweeps with
// a conditional second one here, two modified sweeps there), breakdown below
// sqrt(eps) max(1, ||H[1:j, j]||)
static int arnoldi_build(edk_lindblad *c, int m_target) {
  m_target = std::min(m_target, c->p.m_max);
  const int64_t N = c->N;
  KrylovWork &K = c->K;
  hipStream_t s = c->ctx->stream;
  const double seps = std::sqrt(std::numeric_limits<double>::epsilon());
  for (int j = c->m + 1; j <= m_target; j++) {
    const double2 *vj = c->V + (size_t)(j - 1) * N;
    if (int rc = lv_apply(c->A, vj, c->w)) return rc;
    c->diag.liouvillian_applies++;
    for (int i = 0; i <= c->p.m_max; i++) hred(c, i, j - 1) = 0;
    // A classical Gram-Schmidt sweep, and a second one only when the first
    // removed most of w (Daniel, Gragg, Kaufman and Stewart: ||w'|| <
    // ||w|| / sqrt(2), with ||w||^2 = ||w'||^2 + ||h||^2); both sweeps'
    // coefficients go into the Hessenberg column.  The reference runs two
    // modified sweeps at every step; the entries agree to rounding.
    double hnext = 0;
    for (int sweep = 0; sweep < 2; sweep++) {
      if (int rc = K.cgs_sweep(c->V, N, j, c->w)) return rc;
      EDK_HIP(hipMemcpyAsync(K.h_s, K.d_s, (size_t)j * sizeof(double2), hipMemcpyDeviceToHost, s));
      EDK_HIP(hipStreamSynchronize(s));
      double hh = 0;
      for (int i = 0; i < j; i++) {
        hred(c, i, j - 1) += cplx(K.h_s[i].x, K.h_s[i].y);
        hh += K.h_s[i].x * K.h_s[i].x + K.h_s[i].y * K.h_s[i].y;
      }
      if (int rc = K.norm(c->w, &hnext)) return rc;
      if (hnext * hnext >= 0.5 * (hnext * hnext + hh)) break;
    }
    hred(c, j, j - 1) = hnext;
    c->m = j;
    double cn = 0;
    for (int i = 0; i < j; i++) cn += std::norm(hred(c, i, j - 1));
    if (hnext <= seps * std::max(1.0, std::sqrt(cn))) {
      hred(c, j, j - 1) = 0;
      return EDK_OK;
    }
    if (int rc = K.scale(c->V + (size_t)j * N, c->w, 1.0 / hnext)) return rc;
  }
  return EDK_OK;
}

// _build_reduced! (LindbladEvolution.jl:323-336)
static int build_reduced(edk_lindblad *c) {
  const int m = c->m;
  if (m < 1) return edk_fail(EDK_ERR_NUMERIC, "Cannot build reduced Arnoldi model for empty basis");
  c->Hm.assign((size_t)m * m, 0);
  for (int j = 0; j < m; j++)
    for (int i = 0; i < m; i++) c->Hm[i + (size_t)j * m] = hred(c, i, j);
  c->gamma = onenorm(m, c->Hm.data());
  c->step_cache.clear();
  return EDK_OK;
}

static inline cplx boundary(const edk_lindblad *c) {
  return c->m < 1 ? cplx(0) : c->Hred[c->m + (size_t)(c->m - 1) * (c->p.m_max + 1)];
}

// _reduced_step_matrix (:338-344): exp(tau Hm), cached by the bits of tau
static const std::vector<cplx> &step_matrix(edk_lindblad *c, double tau) {
  uint64_t key;
  memcpy(&key, &tau, sizeof key);
  auto it = c->step_cache.find(key);
  if (it != c->step_cache.end()) return it->second;
  const int m = c->m;
  std::vector<cplx> A((size_t)m * m);
  for (size_t i = 0; i < A.size(); i++) A[i] = c->Hm[i] * tau;
  std::vector<cplx> E;
  cexpm(m, A.data(), E);
  return c->step_cache.emplace(key, std::move(E)).first->second;
}

// _reduced_coeffs_at! (:346-364): exp(tau Hm) e1 in the Arnoldi basis
static void coeffs_at(edk_lindblad *c, double tau, std::vector<cplx> &y) {
  const int m = c->m;
  y.assign(m, 0);
  if (tau == 0) {
    y[0] = 1;
    return;
  }
  const std::vector<cplx> &E = step_matrix(c, tau);
  for (int i = 0; i < m; i++) y[i] = E[i];
}

// _defect_from_y (:366-371)
static double defect_from(const edk_lindblad *c, const std::vector<cplx> &y) {
  const cplx h = boundary(c);
  if (h == cplx(0)) return 0.0;
  return c->norm_anchor * std::abs(h) * std::abs(y[c->m - 1]);
}

// _defect_monitor (:373-376)
static double defect_at(edk_lindblad *c, double tau) {
  std::vector<cplx> y;
  coeffs_at(c, tau, y);
  return defect_from(c, y);
}

// _effective_nsample (:378-385)
static int eff_nsample(const edk_lindblad *c, double tau) {
  const int base = c->p.nsample;
  if (c->m <= 1) return base;
  if (!(c->gamma > 0)) return base;
  const double ne = std::ceil(4.0 * c->gamma * tau) + 1.0;
  return (int)std::min(std::max(ne, (double)base), 513.0);
}

// _max_valid_interval (:387-421)
static double max_valid(edk_lindblad *c, double tau) {
  if (!(tau > 0)) return 0.0;
  if (boundary(c) == cplx(0)) return tau;
  const int n = eff_nsample(c, tau), m = c->m;
  std::vector<cplx> y(m, 0), tmp(m);
  y[0] = 1;
  const double eta0 = defect_from(c, y);
  if (!std::isfinite(eta0) || eta0 > c->p.tol) return 0.0;
  const double step = tau / (double)(n - 1);
  const std::vector<cplx> &S = step_matrix(c, step);
  for (int k = 2; k <= n; k++) {
    for (int i = 0; i < m; i++) {
      cplx a = 0;
      for (int q = 0; q < m; q++) a += S[i + (size_t)q * m] * y[q];
      tmp[i] = a;
    }
    y.swap(tmp);
    const double eta = defect_from(c, y);
    if (!std::isfinite(eta) || eta > c->p.tol) {
      double lo = tau * (double)(k - 2) / (double)(n - 1);
      double hi = (k == n) ? tau : tau * (double)(k - 1) / (double)(n - 1);
      for (int q = 0; q < c->p.bisect_iters; q++) {
        const double mid = 0.5 * (lo + hi);
        if (defect_at(c, mid) > c->p.tol) hi = mid;
        else lo = mid;
      }
      return lo;
    }
  }
  return tau;
}

// _reconstruct_density_raw! (:423-430): out = ||rho_anchor|| V exp(tau Hm) e1
static int reconstruct(edk_lindblad *c, double tau, void *out) {
  std::vector<cplx> y;
  coeffs_at(c, tau, y);
  std::vector<double2> cc(c->m);
  for (int i = 0; i < c->m; i++) cc[i] = make_double2(y[i].real(), y[i].imag());
  return c->K.combine(c->V, c->N, c->m, cc.data(), c->norm_anchor, out, EDK_C64);
}

static int dev_trace(edk_lindblad *c, const void *rho, cplx *tr) {
  hipStream_t s = c->ctx->stream;
  k_lv_trace<<<1, 256, 0, s>>>((const double2 *)rho, c->d, (double2 *)(c->dsc + 2));
  EDK_HIP(hipGetLastError());
  double h[2];
  EDK_HIP(hipMemcpyAsync(h, c->dsc + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  EDK_HIP(hipStreamSynchronize(s));
  *tr = cplx(h[0], h[1]);
  return EDK_OK;
}

// _record_output_diagnostics! (:432-449)
static int record_output(edk_lindblad *c, const void *rho) {
  hipStream_t s = c->ctx->stream;
  const int64_t d = c->d;
  cplx tr;
  if (int rc = dev_trace(c, rho, &tr)) return rc;
  const unsigned g = (unsigned)((d + LV_TILE - 1) / LV_TILE);
  const bool eig = c->p.record_min_eig != 0;
  k_lv_herm<<<dim3(g, g), 256, 0, s>>>((const double2 *)rho, d, c->part, eig ? c->herm : nullptr);
  k_lv_sum<<<1, 256, 0, s>>>(c->part, (int64_t)g * g, c->dsc);
  EDK_HIP(hipGetLastError());
  double h2;
  EDK_HIP(hipMemcpyAsync(&h2, c->dsc, sizeof(double), hipMemcpyDeviceToHost, s));
  EDK_HIP(hipStreamSynchronize(s));
  const double tdrift = std::abs(tr - c->trace_target), hdrift = std::sqrt(h2);
  c->trace_drifts.push_back(tdrift);
  c->herm_drifts.push_back(hdrift);
  c->diag.max_trace_drift = std::max(c->diag.max_trace_drift, tdrift);
  c->diag.max_hermiticity_drift = std::max(c->diag.max_hermiticity_drift, hdrift);
  if (eig) {
    double *ev = c->dsc + 4;
    EDK_RBCHK(rocsolver_zheevd(c->A->h, rocblas_evect_none, rocblas_fill_lower, (rocblas_int)d, rbm(c->herm),
                               (rocblas_int)d, ev, c->ework, c->info));
    double lmin;
    rocblas_int info = 0;
    EDK_HIP(hipMemcpyAsync(&lmin, ev, sizeof(double), hipMemcpyDeviceToHost, s));
    EDK_HIP(hipMemcpyAsync(&info, c->info, sizeof(rocblas_int), hipMemcpyDeviceToHost, s));
    EDK_HIP(hipStreamSynchronize(s));
    if (info != 0) return edk_fail(EDK_ERR_NUMERIC, "record_min_eig: eigensolver did not converge");
    c->min_eigs.push_back(lmin);
    c->diag.min_min_hermitian_eig = std::min(c->diag.min_min_hermitian_eig, lmin);
  }
  return EDK_OK;
}

// _extend_basis! (:460-473)
static int extend(edk_lindblad *c, int *did) {
  *did = 0;
  if (!c->p.extend_basis || c->m >= c->p.m_max) return EDK_OK;
  const int old = c->m;
  if (int rc = arnoldi_build(c, std::min(c->m + c->p.extend_step, c->p.m_max))) return rc;
  if (c->m > old) {
    if (int rc = build_reduced(c)) return rc;
    c->diag.basis_extensions++;
    c->diag.max_dim_used = std::max<int64_t>(c->diag.max_dim_used, c->m);
    *did = 1;
  }
  return EDK_OK;
}

// _restart_from_new_anchor! (:475-494)
static int restart(edk_lindblad *c, double tau) {
  if (int rc = reconstruct(c, tau, c->rho_anchor)) return rc;
  c->t_anchor += tau;
  double nrm;
  if (int rc = c->K.norm(c->rho_anchor, &nrm)) return rc;
  if (!(nrm > 0))
    return edk_fail(EDK_ERR_NUMERIC,
                    "LindbladArnoldiCache: restart failed because anchor density matrix has zero Frobenius norm");
  c->norm_anchor = nrm;
  std::fill(c->Hred.begin(), c->Hred.end(), cplx(0));
  c->m = 0;
  if (int rc = c->K.scale(c->V, c->rho_anchor, 1.0 / nrm)) return rc;
  if (int rc = arnoldi_build(c, c->p.m_init)) return rc;
  if (int rc = build_reduced(c)) return rc;
  c->diag.basis_builds++;
  c->diag.restarts++;
  c->diag.max_dim_used = std::max<int64_t>(c->diag.max_dim_used, c->m);
  return EDK_OK;
}

// _ensure_valid_up_to! (:496-535); the extension branch for tau_valid <= 0
// is the same call as the one before it and is folded into it
static int ensure_valid(edk_lindblad *c, double target) {
  double rem = target;
  if (rem <= 0) return EDK_OK;
  while (true) {
    const double v = max_valid(c, rem);
    if (v >= rem) return EDK_OK;
    int did = 0;
    if (c->m < c->p.m_max) {
      if (int rc = extend(c, &did)) return rc;
      if (did) continue;
    }
    if (v <= 0)
      return edk_fail(EDK_ERR_NUMERIC,
                      "LindbladArnoldiCache: basis is not valid at tau=0; consider increasing m_max or loosening tol");
    c->accepted.push_back(v);
    c->diag.n_accepted_intervals++;
    if (int rc = restart(c, v)) return rc;
    rem -= v;
  }
}

static void lb_free(edk_lindblad *c) {
  if (!c) return;
  edk_dfree(c->ctx, c->rho_anchor);
  edk_dfree(c->ctx, c->V);
  edk_dfree(c->ctx, c->w);
  edk_dfree(c->ctx, c->part);
  edk_dfree(c->ctx, c->dsc);
  edk_dfree(c->ctx, c->herm);
  edk_dfree(c->ctx, c->ework);
  edk_dfree(c->ctx, c->info);
  c->K.release();
  delete c;
}

// LindbladArnoldiCache(A, dm; kwargs...) (LindbladEvolution.jl:216-273);
// rho0: device d x d complex128, column-major
extern "C" int edk_lindblad_create(edk_liouvillian *A, const void *rho0, const edk_lindblad_params *params,
                                   edk_lindblad **out) {
  if (!A || !rho0 || !out) return edk_fail(EDK_ERR_ARG, "edk_lindblad_create: null argument");
  edk_lindblad_params p;
  edk_lindblad_default_params(&p);
  if (params) p = *params;
  if (p.m_init < 1) return edk_fail(EDK_ERR_ARG, "m_init must be >= 1");
  if (p.m_max < p.m_init) return edk_fail(EDK_ERR_ARG, "m_max must be >= m_init");
  if (p.extend_step < 1) return edk_fail(EDK_ERR_ARG, "extend_step must be >= 1");
  if (p.nsample < 2) return edk_fail(EDK_ERR_ARG, "nsample must be >= 2");
  if (p.bisect_iters < 0) return edk_fail(EDK_ERR_ARG, "bisect_iters must be >= 0");
  const int64_t d = A->d, N = d * d;
  if ((int64_t)p.m_max > N) p.m_max = (int)N;
  if (p.m_init > p.m_max) p.m_init = p.m_max;
  if (p.m_max > EDK_KRYLOV_MAXM) return edk_fail(EDK_ERR_UNSUPPORTED, "m_max above 1024");
  edk_ctx *ctx = A->ctx;
  EDK_HIP(hipSetDevice(ctx->device));
  edk_lindblad *c = new edk_lindblad();
  c->A = A;
  c->ctx = ctx;
  c->d = d;
  c->N = N;
  c->p = p;
  memset(&c->diag, 0, sizeof(c->diag));
  c->diag.min_min_hermitian_eig = INFINITY;
  c->Hred.assign((size_t)(p.m_max + 1) * p.m_max, cplx(0));
  const unsigned g = (unsigned)((d + LV_TILE - 1) / LV_TILE);
  c->nparts = (int64_t)g * g;
  hipStream_t s = ctx->stream;
  int rc = c->K.init(ctx, N, EDK_C64, p.m_max + 1);
  const size_t vb = (size_t)N * sizeof(double2);
  if (!rc && (edk_dalloc(ctx, (void **)&c->rho_anchor, vb) != hipSuccess ||
              edk_dalloc(ctx, (void **)&c->V, vb * (p.m_max + 1)) != hipSuccess ||
              edk_dalloc(ctx, (void **)&c->w, vb) != hipSuccess ||
              edk_dalloc(ctx, (void **)&c->part, (size_t)c->nparts * sizeof(double)) != hipSuccess ||
              edk_dalloc(ctx, (void **)&c->dsc, (size_t)(d + 8) * sizeof(double)) != hipSuccess))
    rc = edk_fail(EDK_ERR_OOM, "LindbladArnoldiCache does not fit in device memory");
  if (!rc && p.record_min_eig &&
      (edk_dalloc(ctx, (void **)&c->herm, vb) != hipSuccess ||
       edk_dalloc(ctx, (void **)&c->ework, (size_t)d * sizeof(double)) != hipSuccess ||
       edk_dalloc(ctx, (void **)&c->info, sizeof(rocblas_int)) != hipSuccess))
    rc = edk_fail(EDK_ERR_OOM, "LindbladArnoldiCache does not fit in device memory");
  if (!rc && hipMemcpyAsync(c->rho_anchor, rho0, vb, hipMemcpyDeviceToDevice, s) != hipSuccess)
    rc = edk_fail(EDK_ERR_HIP, "copy rho0");
  double nrm = 0;
  if (!rc) rc = c->K.norm(c->rho_anchor, &nrm);
  if (!rc && !(nrm > 0)) rc = edk_fail(EDK_ERR_NUMERIC, "LindbladArnoldiCache: initial density matrix has zero Frobenius norm");
  if (!rc) rc = dev_trace(c, c->rho_anchor, &c->trace_target);
  if (!rc) {
    c->norm_anchor = nrm;
    rc = c->K.scale(c->V, c->rho_anchor, 1.0 / nrm);
  }
  if (!rc) rc = arnoldi_build(c, p.m_init);
  if (!rc) rc = build_reduced(c);
  if (rc) {
    lb_free(c);
    return rc;
  }
  c->diag.basis_builds = 1;
  c->diag.max_dim_used = c->m;
  *out = c;
  return EDK_OK;
}

// lindblad_timeevolve!(cache, t) (LindbladEvolution.jl:598-611); out: device d x d
extern "C" int edk_lindblad_evolve(edk_lindblad *c, double t, void *out) {
  if (!c || !out) return edk_fail(EDK_ERR_ARG, "edk_lindblad_evolve: null argument");
  EDK_HIP(hipSetDevice(c->ctx->device));
  if (t < c->t_latest)
    return edk_fail(EDK_ERR_ARG,
                    "lindblad_timeevolve!: requested time precedes the last served time; backward evolution is not supported");
  const double tau = t - c->t_anchor;
  if (tau < 0) return edk_fail(EDK_ERR_ARG, "lindblad_timeevolve!: requested time precedes anchor time");
  if (int rc = ensure_valid(c, tau)) return rc;
  if (int rc = reconstruct(c, t - c->t_anchor, out)) return rc;
  if (int rc = record_output(c, out)) return rc;
  c->t_latest = t;
  c->diag.total_times_served++;
  // _copy_output_density (:451-458)
  if (c->p.normalize_trace) {
    cplx tr;
    if (int rc = dev_trace(c, out, &tr)) return rc;
    if (tr != cplx(0)) {
      const cplx f = c->trace_target / tr;
      k_lv_cscale<<<(unsigned)((c->N + 255) / 256), 256, 0, c->ctx->stream>>>((double2 *)out,
                                                                               make_double2(f.real(), f.imag()), c->N);
      EDK_HIP(hipGetLastError());
    }
  }
  EDK_HIP(hipStreamSynchronize(c->ctx->stream));
  return EDK_OK;
}

extern "C" int edk_lindblad_diagnostics(edk_lindblad *c, edk_lindblad_diag *d) {
  if (!c || !d) return edk_fail(EDK_ERR_ARG, "null argument");
  *d = c->diag;
  return EDK_OK;
}

extern "C" int edk_lindblad_records(edk_lindblad *c, double *accepted, double *trace_drifts, double *herm_drifts,
                                    double *min_eigs, int64_t cap, int64_t *n_accepted, int64_t *n_outputs,
                                    int64_t *n_eigs) {
  if (!c || !n_accepted || !n_outputs || !n_eigs) return edk_fail(EDK_ERR_ARG, "null argument");
  *n_accepted = (int64_t)c->accepted.size();
  *n_outputs = (int64_t)c->trace_drifts.size();
  *n_eigs = (int64_t)c->min_eigs.size();
  for (int64_t i = 0; i < std::min<int64_t>(cap, *n_accepted); i++)
    if (accepted) accepted[i] = c->accepted[i];
  for (int64_t i = 0; i < std::min<int64_t>(cap, *n_outputs); i++) {
    if (trace_drifts) trace_drifts[i] = c->trace_drifts[i];
    if (herm_drifts) herm_drifts[i] = c->herm_drifts[i];
  }
  for (int64_t i = 0; i < std::min<int64_t>(cap, *n_eigs); i++)
    if (min_eigs) min_eigs[i] = c->min_eigs[i];
  return EDK_OK;
}

extern "C" int edk_lindblad_dim(edk_lindblad *c, int *m) {
  if (!c || !m) return edk_fail(EDK_ERR_ARG, "null argument");
  *m = c->m;
  return EDK_OK;
}

extern "C" int edk_lindblad_destroy(edk_lindblad *c) {
  if (!c) return EDK_OK;
  hipSetDevice(c->ctx->device);
  hipStreamSynchronize(c->ctx->stream);
  lb_free(c);
  return EDK_OK;
}
