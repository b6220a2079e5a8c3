// edk_schmidt.hip -- the Schmidt matrix of a sector state and its spectrum
// on the device: schmidt(v, Ainds, b) (Schmidt.jl:186-330,
// AbelianBasis.jl:704-719) and the squared Schmidt values behind ent_S
// (Schmidt.jl:143-160).
//
// One lane per representative walks the representative's orbit in the
// reference's order, with the reference's running value (v[i] / R[i], times
// the symmetry factor of each step), and adds it at
// M[index(A digits), index(B digits)] (schmidtmatrix / addto!,
// Schmidt.jl:39-72).  Distinct representatives have disjoint orbits and a
// matrix entry is one product state, so every entry is written by a single
// lane, in the reference's order: no atomics, and the result does not depend
// on scheduling.
//
// The spectrum: rho = M M^H (or M^H M, whichever is smaller) with rocBLAS
// zherk, then its eigenvalues with rocSOLVER zheevd; they are the squared
// singular values of M that ent_S takes (svdvals(...) .^ 2).
#include <climits>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <vector>

#include "edk_matvec.h"

#define EDK_SCHMIDT_MAXL 64

struct SchmidtArgs {
  CanonDev C;
  int kind, L, base, w;
  int nA, nB;
  int shA[EDK_SCHMIDT_MAXL], shB[EDK_SCHMIDT_MAXL];  // bit offset of each A / B site's field
  int64_t dA;
  double2 step;  // translation factor per shift (b.C[2])
};

__device__ __forceinline__ void schmidt_add(const SchmidtArgs &a, double2 *M, uint64_t s, double2 val) {
  const uint64_t fm = (1ull << a.w) - 1;
  int64_t ia = 0, ib = 0;
  for (int q = 0; q < a.nA; q++) ia = ia * a.base + (int64_t)((s >> a.shA[q]) & fm);
  for (int q = 0; q < a.nB; q++) ib = ib * a.base + (int64_t)((s >> a.shB[q]) & fm);
  double2 &m = M[ia + a.dA * ib];
  m.x += val.x;
  m.y += val.y;
}

__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

template <typename VT>
__global__ void k_schmidt(SchmidtArgs a, const uint64_t *keys, const double *R, const VT *v, int64_t D,
                          double2 *M) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D) return;
  const CanonDev &C = a.C;
  const VT vi = v[i];
  if (a.kind == EDK_TENSOR || a.kind == EDK_PROJECTED) {  // AbstractOnsiteBasis (Schmidt.jl:186-195)
    const uint64_t s = a.kind == EDK_TENSOR ? edk_content_to_packed_d((uint64_t)i, a.L, a.base, a.w) : keys[i];
    schmidt_add(a, M, s, make_double2((double)vi.x, (double)vi.y));
    return;
  }
  uint64_t s = keys[i];
  const double r = R[i];
  double2 val = make_double2((double)vi.x / r, (double)vi.y / r);
  const int tot = C.L * C.w, sh = C.A * C.w;
  if (a.kind == EDK_TRANSLATIONAL || a.kind == EDK_TRANSLATION_PARITY || a.kind == EDK_TRANSLATION_FLIP) {
    // Schmidt.jl:202-216 and parity_schmidt :249-273: len shifts by the unit
    // cell (circshift!), the value times b.C[2] after each; then the reflected
    // (or flipped) word, the value times P, and len shifts again
    for (int j = 0; j < C.len; j++) {
      schmidt_add(a, M, s, val);
      s = C.len > 1 ? edk_rotr(s, sh, tot, C.mask) : s;
      val = cmul(val, a.step);
    }
    if (a.kind == EDK_TRANSLATIONAL) return;
    s = a.kind == EDK_TRANSLATION_PARITY ? edk_revfields(s, C.L, C.w) : edk_flipfields(C, s);
    val = cscale(val, (double)C.P);
    for (int j = 0; j < C.len; j++) {
      schmidt_add(a, M, s, val);
      s = C.len > 1 ? edk_rotr(s, sh, tot, C.mask) : s;
      val = cmul(val, a.step);
    }
    return;
  }
  if (a.kind == EDK_PARITY || a.kind == EDK_FLIP) {  // Schmidt.jl:279-309
    schmidt_add(a, M, s, val);
    s = a.kind == EDK_PARITY ? edk_revfields(s, C.L, C.w) : edk_flipfields(C, s);
    schmidt_add(a, M, s, cscale(val, (double)C.P));
    return;
  }
  if (a.kind == EDK_PARITY_FLIP) {  // Schmidt.jl:313-330: 1, P, PZ, Z
    schmidt_add(a, M, s, val);
    s = edk_revfields(s, C.L, C.w);
    val = cscale(val, (double)C.P);
    schmidt_add(a, M, s, val);
    s = edk_flipfields(C, s);
    val = cscale(val, (double)C.Z);
    schmidt_add(a, M, s, val);
    s = edk_revfields(s, C.L, C.w);
    val = cscale(val, (double)C.P);
    schmidt_add(a, M, s, val);
    return;
  }
  // AbelianBasis (AbelianBasis.jl:704-719): every group element in odometer
  // order, the value times phase(g)
  int os[EDK_MAXGEN];
  for (int g = 0; g < C.ngen; g++) os[g] = 1;
  schmidt_add(a, M, s, val);
  for (int step = 2; step <= C.order; step++) {
    for (int g = 0; g < C.ngen; g++) {
      s = edk_apply_gen(C, g, s);
      if (++os[g] > C.gdesc[EDK_GDESC * g + 1]) os[g] = 1;
      else break;
    }
    double2 ph = C.ph[C.gdesc[6] + os[0] - 1];
    for (int g = 1; g < C.ngen; g++) ph = cmul(ph, C.ph[C.gdesc[EDK_GDESC * g + 6] + os[g] - 1]);
    schmidt_add(a, M, s, cmul(ph, val));
  }
}

__global__ void k_zero_c(double2 *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_double2(0.0, 0.0);
}

static int schmidt_setup(const edk_basis *b, int nA, const int32_t *Ainds, SchmidtArgs &a) {
  if (b->L > EDK_SCHMIDT_MAXL) return edk_fail(EDK_ERR_UNSUPPORTED, "schmidt: L above 64");
  if (nA < 0 || nA > b->L) return edk_fail(EDK_ERR_ARG, "schmidt: bad subsystem size");
  std::vector<char> inA(b->L + 1, 0);
  for (int q = 0; q < nA; q++) {
    if (Ainds[q] < 1 || Ainds[q] > b->L || inA[Ainds[q]])
      return edk_fail(EDK_ERR_ARG, "schmidt: subsystem sites must be distinct and in 1:L");
    inA[Ainds[q]] = 1;
  }
  a.C = b->canon();
  a.kind = b->kind;
  a.L = b->L;
  a.base = b->base;
  a.w = b->w;
  a.nA = nA;
  a.nB = b->L - nA;
  // site q (1-based) is the field at bit (L - q) * w (site 1 most significant)
  for (int q = 0; q < nA; q++) a.shA[q] = (b->L - Ainds[q]) * b->w;
  int nb = 0;
  for (int q = 1; q <= b->L; q++)  // Binds: the other sites in increasing order (Schmidt.jl:46-52)
    if (!inA[q]) a.shB[nb++] = (b->L - q) * b->w;
  double dA = 1, dB = 1;
  for (int q = 0; q < a.nA; q++) dA *= b->base;
  for (int q = 0; q < a.nB; q++) dB *= b->base;
  if (dA * dB > (double)(1ll << 34)) return edk_fail(EDK_ERR_UNSUPPORTED, "schmidt: matrix above 2^34 entries");
  a.dA = (int64_t)dA;
  // b.C[2]: the translation factor per shift (TranslationParityBasis: +-1)
  a.step = make_double2(1.0, 0.0);
  if (b->kind == EDK_TRANSLATIONAL || b->kind == EDK_TRANSLATION_FLIP) {
    if (b->len > 1) a.step = b->h_ph[1];
  } else if (b->kind == EDK_TRANSLATION_PARITY) {
    a.step = make_double2(b->kp == 1 ? 1.0 : -1.0, 0.0);
  }
  return EDK_OK;
}

static int schmidt_build(edk_basis *b, int dtype, const void *v, const SchmidtArgs &a, double2 *M, int64_t n) {
  edk_ctx *ctx = b->ctx;
  const unsigned gz = (unsigned)((n + 255) / 256);
  if (n > 0) k_zero_c<<<gz, 256, 0, ctx->stream>>>(M, n);
  const int64_t D = b->D;
  if (D > 0) {
    const unsigned g = (unsigned)((D + 255) / 256);
    if (dtype == EDK_C64)
      k_schmidt<double2><<<g, 256, 0, ctx->stream>>>(a, b->d_keys, b->d_R, (const double2 *)v, D, M);
    else
      k_schmidt<float2><<<g, 256, 0, ctx->stream>>>(a, b->d_keys, b->d_R, (const float2 *)v, D, M);
  }
  EDK_HIP(hipGetLastError());
  return EDK_OK;
}

extern "C" int edk_schmidt(edk_basis *b, int dtype, const void *v, int nA, const int32_t *Ainds, void *M) {
  if (!b || !v || !M || (nA > 0 && !Ainds)) return edk_fail(EDK_ERR_ARG, "edk_schmidt: null argument");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "edk_schmidt: bad dtype");
  if (b->uploaded && b->kind != EDK_TENSOR && b->kind != EDK_PROJECTED)
    return edk_fail(EDK_ERR_UNSUPPORTED, "edk_schmidt: needs a device-built basis (orbit tables)");
  SchmidtArgs a;
  if (int rc = schmidt_setup(b, nA, Ainds, a)) return rc;
  EDK_HIP(hipSetDevice(b->ctx->device));
  int64_t dB = 1;
  for (int q = 0; q < a.nB; q++) dB *= b->base;
  if (int rc = schmidt_build(b, dtype, v, a, (double2 *)M, a.dA * dB)) return rc;
  EDK_HIP(hipStreamSynchronize(b->ctx->stream));
  return EDK_OK;
}

#define EDK_RB(call)                                                                  \
  do {                                                                                \
    rocblas_status st_ = (call);                                                      \
    if (st_ != rocblas_status_success) {                                              \
      rc = edk_fail(EDK_ERR_HIP, std::string("rocBLAS/rocSOLVER: ") + rocblas_status_to_string(st_)); \
      goto done;                                                                      \
    }                                                                                 \
  } while (0)

extern "C" int edk_ent_spec(edk_basis *b, int dtype, const void *v, int nA, const int32_t *Ainds, double *s2,
                            int64_t *n_out) {
  if (!b || !v || !s2 || !n_out || (nA > 0 && !Ainds)) return edk_fail(EDK_ERR_ARG, "edk_ent_spec: null argument");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "edk_ent_spec: bad dtype");
  if (b->uploaded && b->kind != EDK_TENSOR && b->kind != EDK_PROJECTED)
    return edk_fail(EDK_ERR_UNSUPPORTED, "edk_ent_spec: needs a device-built basis (orbit tables)");
  SchmidtArgs a;
  if (int rc = schmidt_setup(b, nA, Ainds, a)) return rc;
  edk_ctx *ctx = b->ctx;
  EDK_HIP(hipSetDevice(ctx->device));
  int64_t dB = 1;
  for (int q = 0; q < a.nB; q++) dB *= b->base;
  const int64_t dA = a.dA, n = std::min(dA, dB), k = std::max(dA, dB);
  if (n > 32768) return edk_fail(EDK_ERR_UNSUPPORTED, "edk_ent_spec: reduced density matrix above 32768^2");
  // rocBLAS takes the inner size k and the leading dimension dA as 32-bit
  // integers: reject before anything is allocated or built
  if (k > (int64_t)INT32_MAX || dA > (int64_t)INT32_MAX)
    return edk_fail(EDK_ERR_UNSUPPORTED, "edk_ent_spec: Schmidt matrix dimension above 2^31 - 1");
  int rc = EDK_OK;
  double2 *M = nullptr, *rho = nullptr;
  double *d_w = nullptr, *d_e = nullptr;
  rocblas_int *d_info = nullptr;
  rocblas_handle h = nullptr;
  rocblas_int info = 0;
  std::vector<double> w(n);
  const double one = 1.0, zero = 0.0;
  if (hipMalloc((void **)&M, (size_t)dA * dB * sizeof(double2)) != hipSuccess ||
      hipMalloc((void **)&rho, (size_t)n * n * sizeof(double2)) != hipSuccess ||
      hipMalloc((void **)&d_w, n * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&d_e, n * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&d_info, sizeof(rocblas_int)) != hipSuccess) {
    rc = edk_fail(EDK_ERR_OOM, "edk_ent_spec: Schmidt matrix does not fit in device memory");
    goto done;
  }
  if ((rc = schmidt_build(b, dtype, v, a, M, dA * dB))) goto done;
  EDK_RB(rocblas_create_handle(&h));
  EDK_RB(rocblas_set_stream(h, ctx->stream));
  // rho = M M^H (dA <= dB) or M^H M, lower triangle
  EDK_RB(rocblas_zherk(h, rocblas_fill_lower, dA <= dB ? rocblas_operation_none : rocblas_operation_conjugate_transpose,
                       (rocblas_int)n, (rocblas_int)k, &one, (const rocblas_double_complex *)M, (rocblas_int)dA, &zero,
                       (rocblas_double_complex *)rho, (rocblas_int)n));
  EDK_RB(rocsolver_zheevd(h, rocblas_evect_none, rocblas_fill_lower, (rocblas_int)n, (rocblas_double_complex *)rho,
                          (rocblas_int)n, d_w, d_e, d_info));
  if (hipMemcpyAsync(w.data(), d_w, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipMemcpyAsync(&info, d_info, sizeof(rocblas_int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    rc = edk_fail(EDK_ERR_HIP, "edk_ent_spec: copy of the spectrum");
    goto done;
  }
  if (info != 0) {
    rc = edk_fail(EDK_ERR_NUMERIC, "edk_ent_spec: eigensolver did not converge");
    goto done;
  }
  // descending, as svdvals(...) .^ 2; rounding below zero clamped
  for (int64_t q = 0; q < n; q++) s2[q] = std::max(w[n - 1 - q], 0.0);
  *n_out = n;
done:
  if (h) (void)rocblas_destroy_handle(h);
  (void)hipFree(M);
  (void)hipFree(rho);
  (void)hipFree(d_w);
  (void)hipFree(d_e);
  (void)hipFree(d_info);
  return rc;
}
