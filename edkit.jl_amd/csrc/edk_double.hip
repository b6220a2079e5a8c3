// edk_double.hip -- DoubleBasis (LinearMap.jl:88-185) on the device: the
// rectangular operator Operator(M, I, DoubleBasis(B1, B2)) from the source
// sector B2 to the target sector B1, and the basis-overlap map (B::DoubleBasis)(v).
//
// Both are row gathers over the target sector: one lane per B1
// representative a_i, no atomics.
//
// Operator.  The reference scatters column j (Operator.jl:418-459 with
// change!/index of DoubleBasis, LinearMap.jl:115-126): for the B2
// representative a_j, every term t and every nonzero M_t[row, loc_t(a_j)],
//   H[i, j] += M_t[row, loc_t(a_j)] * coeff1(b') / R2_j,   b' = a_j with S_t <- row,
// where (coeff1, i) = index(B1, b').  The pairs (j, row) and (b', c) with
// b' in the B1 orbit of a_i, c = loc_t(a_j) and a_j = b' with S_t <- c are in
// one-to-one correspondence, so row i is
//   y_i = sum over distinct b' in orbit1(a_i) of coeff1(b') *
//         sum_t sum_c M_t[loc_t(b'), c] x_j / R2_j   (a_j = b' with S_t <- c a B2 representative).
// The lane walks the group of B1 (every element, so each distinct orbit state
// |stab(a_i)| times) and divides by the number of elements that fix a_i.
//
// Map.  (B::DoubleBasis)(v) sums over the full tensor space
// (LinearMap.jl:144-172): out[pos1] += conj(coeff1/ord1) (coeff2/ord2) v[pos2];
// the states with pos1 = i are the distinct orbit states of a_i, walked the
// same way.
#include <algorithm>
#include <vector>

#include "edk_matvec.h"

struct SideArgs {
  CanonDev C;
  SearchDev S;
  const double *R;
  const uint64_t *keys;
  int kind;
  int64_t D;
};

struct DoubleArgs {
  SideArgs b1, b2;  // target, source
  TermsDev T;       // local terms (operator)
  int L, base, w;
  double inv_ord;   // 1 / (orbit_order(B1) orbit_order(B2)) (map)
};

__device__ __forceinline__ uint64_t dbl_content(uint64_t s, int L, int base, int w) {
  if (base == 2) return s;
  const uint64_t fm = (1ull << w) - 1;
  uint64_t c = 0;
  for (int q = L - 1; q >= 0; q--) c = c * (uint64_t)base + ((s >> (q * w)) & fm);
  return c;
}

// index_nocheck(B, s) (LinearMap.jl:13-16): position (-1 when absent) and coefficient
__device__ __forceinline__ int64_t side_index(const SideArgs &B, uint64_t s, int L, int base, int w,
                                              double2 &coef) {
  coef = make_double2(1.0, 0.0);
  if (B.kind == EDK_TENSOR) return (int64_t)dbl_content(s, L, base, w);
  if (B.kind == EDK_PROJECTED) return edk_search(B.S, s);
  const CanonOut o = edk_canon(B.C, s);
  const int64_t j = edk_search(B.S, o.rep);
  if (j < 0) return -1;
  const double r = B.R[j];
  coef = make_double2(o.phase.x * r, o.phase.y * r);
  return j;
}

// position of s as a representative of B (-1: not one)
__device__ __forceinline__ int64_t side_rep(const SideArgs &B, uint64_t s, int L, int base, int w) {
  if (B.kind == EDK_TENSOR) return (int64_t)dbl_content(s, L, base, w);
  return edk_search(B.S, s);
}

// norm(B, j): the value change! returns (1 for on-site bases)
__device__ __forceinline__ double side_norm(const SideArgs &B, int64_t j) {
  return (B.kind == EDK_TENSOR || B.kind == EDK_PROJECTED) ? 1.0 : B.R[j];
}

__device__ __forceinline__ uint64_t side_state(const SideArgs &B, int64_t i, int L, int base, int w) {
  return B.kind == EDK_TENSOR ? edk_content_to_packed_d((uint64_t)i, L, base, w) : B.keys[i];
}

// f(g s) for every element g of the basis' group, in the order of the
// per-kind orbit walks (translations by the unit cell, then the reflected or
// flipped word; 1, P, PZ, Z; odometer order over Abelian generators)
template <typename F>
__device__ __forceinline__ void walk_group(const CanonDev &C, int kind, uint64_t s, F &&f) {
  if (kind == EDK_TENSOR || kind == EDK_PROJECTED) {
    f(s);
    return;
  }
  const int tot = C.L * C.w, sh = C.A * C.w;
  if (kind == EDK_TRANSLATIONAL || kind == EDK_TRANSLATION_PARITY || kind == EDK_TRANSLATION_FLIP) {
    for (int j = 0; j < C.len; j++) {
      f(s);
      s = C.len > 1 ? edk_rotr(s, sh, tot, C.mask) : s;
    }
    if (kind == EDK_TRANSLATIONAL) return;
    s = kind == EDK_TRANSLATION_PARITY ? edk_revfields(s, C.L, C.w) : edk_flipfields(C, s);
    for (int j = 0; j < C.len; j++) {
      f(s);
      s = C.len > 1 ? edk_rotr(s, sh, tot, C.mask) : s;
    }
    return;
  }
  if (kind == EDK_PARITY || kind == EDK_FLIP) {
    f(s);
    f(kind == EDK_PARITY ? edk_revfields(s, C.L, C.w) : edk_flipfields(C, s));
    return;
  }
  if (kind == EDK_PARITY_FLIP) {
    f(s);
    s = edk_revfields(s, C.L, C.w);
    f(s);
    s = edk_flipfields(C, s);
    f(s);
    s = edk_revfields(s, C.L, C.w);
    f(s);
    return;
  }
  int os[EDK_MAXGEN];
  for (int g = 0; g < C.ngen; g++) os[g] = 1;
  f(s);
  for (int step = 2; step <= C.order; step++) {
    for (int g = 0; g < C.ngen; g++) {
      s = edk_apply_gen(C, g, s);
      if (++os[g] > C.gdesc[EDK_GDESC * g + 1]) os[g] = 1;
      else break;
    }
    f(s);
  }
}

template <typename VT, typename ST, int NV>
__global__ void __launch_bounds__(256) k_double_op(DoubleArgs A, const VT *__restrict__ x, VT *__restrict__ y,
                                                   int64_t nb, int64_t col, int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.b1.D) return;
  const TermsDev &T = A.T;
  const int L = A.L, base = A.base, w = A.w;
  const uint64_t fm = (1ull << w) - 1;
  const uint64_t s0 = side_state(A.b1, i, L, base, w);
  ST ar[NV], ai[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) ar[c] = ai[c] = 0;
  int stab = 0;
  walk_group(A.b1.C, A.b1.kind, s0, [&](uint64_t bp) {
    stab += bp == s0;
    double2 c1;
    if (side_index(A.b1, bp, L, base, w, c1) < 0) return;
    ST pr[NV], pi[NV];
#pragma unroll
    for (int c = 0; c < NV; c++) pr[c] = pi[c] = 0;
    for (int t = 0; t < T.nterms; t++) {
      const int *td = T.tdesc + t * EDK_TDESC;
      const int ns = td[0], m = td[1];
      int loc = 0;
      for (int q = 0; q < ns; q++) loc = loc * base + (int)((bp >> td[2 + q]) & fm);
      const int *rp = T.rowptr + T.mrow_off[m];
      for (int e = rp[loc]; e < rp[loc + 1]; e++) {
        int cc = T.col[e];
        uint64_t a = bp;
        for (int q = ns - 1; q >= 0; q--) {
          const uint64_t nd = (uint64_t)(cc % base);
          cc /= base;
          const int sh = td[2 + q];
          a = (a & ~(fm << sh)) | (nd << sh);
        }
        const int64_t j = side_rep(A.b2, a, L, base, w);
        if (j < 0) continue;
        const double rn = 1.0 / side_norm(A.b2, j);
        const double2 v = T.val[e];
        const ST vr = (ST)(v.x * rn), vi = (ST)(v.y * rn);
        const VT *xj = x + j * nb + col;
#pragma unroll
        for (int c = 0; c < NV; c++) {
          const VT xv = xj[c];
          pr[c] += vr * (ST)xv.x - vi * (ST)xv.y;
          pi[c] += vr * (ST)xv.y + vi * (ST)xv.x;
        }
      }
    }
    const ST cr = (ST)c1.x, ci = (ST)c1.y;
#pragma unroll
    for (int c = 0; c < NV; c++) {
      ar[c] += cr * pr[c] - ci * pi[c];
      ai[c] += cr * pi[c] + ci * pr[c];
    }
  });
  const ST inv = (ST)1 / (ST)stab;
#pragma unroll
  for (int c = 0; c < NV; c++) {
    VT out;
    out.x = ar[c] * inv;
    out.y = ai[c] * inv;
    if (accumulate) {
      const VT y0 = y[i * nb + col + c];
      out.x += y0.x;
      out.y += y0.y;
    }
    y[i * nb + col + c] = out;
  }
}

template <typename VT, typename ST, int NV>
__global__ void __launch_bounds__(256) k_double_map(DoubleArgs A, const VT *__restrict__ x, VT *__restrict__ y,
                                                    int64_t nb, int64_t col) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.b1.D) return;
  const int L = A.L, base = A.base, w = A.w;
  const uint64_t s0 = side_state(A.b1, i, L, base, w);
  ST ar[NV], ai[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) ar[c] = ai[c] = 0;
  int stab = 0;
  walk_group(A.b1.C, A.b1.kind, s0, [&](uint64_t s) {
    stab += s == s0;
    double2 c1, c2;
    if (side_index(A.b1, s, L, base, w, c1) < 0) return;
    const int64_t j = side_index(A.b2, s, L, base, w, c2);
    if (j < 0) return;
    // conj(coeff1 / ord1) * (coeff2 / ord2)
    const ST kr = (ST)((c1.x * c2.x + c1.y * c2.y) * A.inv_ord), ki = (ST)((c1.x * c2.y - c1.y * c2.x) * A.inv_ord);
    const VT *xj = x + j * nb + col;
#pragma unroll
    for (int c = 0; c < NV; c++) {
      const VT xv = xj[c];
      ar[c] += kr * (ST)xv.x - ki * (ST)xv.y;
      ai[c] += kr * (ST)xv.y + ki * (ST)xv.x;
    }
  });
  const ST inv = (ST)1 / (ST)stab;
#pragma unroll
  for (int c = 0; c < NV; c++) {
    VT out;
    out.x = ar[c] * inv;
    out.y = ai[c] * inv;
    y[i * nb + col + c] = out;
  }
}

template <typename VT>
__global__ void k_double_zero(VT *y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    VT z;
    z.x = 0;
    z.y = 0;
    y[i] = z;
  }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
// an empty source sector: the map is the zero matrix, no lookups
static void zero_out(int dtype, void *y, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  const unsigned g = (unsigned)((n + 255) / 256);
  if (dtype == EDK_C64) k_double_zero<double2><<<g, 256, 0, st>>>((double2 *)y, n);
  else k_double_zero<float2><<<g, 256, 0, st>>>((float2 *)y, n);
}

struct edk_dop {
  edk_op *op;          // the terms, on B1
  edk_basis *b2;
};

static SideArgs side_of(const edk_basis *b) {
  SideArgs s;
  s.C = b->canon();
  s.S = b->search();
  s.R = b->d_R;
  s.keys = b->d_keys;
  s.kind = b->kind;
  s.D = b->D;
  return s;
}

static int orbit_order(const edk_basis *b) {
  return (b->kind == EDK_TENSOR || b->kind == EDK_PROJECTED) ? 1 : b->order;
}

// check_doublebasis_compatible (LinearMap.jl:39-43) and the device tables
static int double_setup(const edk_basis *b1, const edk_basis *b2, DoubleArgs &a) {
  if (!b1 || !b2) return edk_fail(EDK_ERR_ARG, "DoubleBasis: null basis");
  if (b1->L != b2->L) return edk_fail(EDK_ERR_ARG, "DoubleBasis requires bases with the same length.");
  if (b1->base != b2->base) return edk_fail(EDK_ERR_ARG, "DoubleBasis requires bases with the same local base.");
  if (b1->ctx->device != b2->ctx->device) return edk_fail(EDK_ERR_ARG, "DoubleBasis: bases on different devices");
  for (const edk_basis *b : {b1, b2})
    if (b->uploaded && b->kind != EDK_TENSOR && b->kind != EDK_PROJECTED)
      return edk_fail(EDK_ERR_UNSUPPORTED, "DoubleBasis: needs device-built bases (orbit tables)");
  a.b1 = side_of(b1);
  a.b2 = side_of(b2);
  a.L = b1->L;
  a.base = b1->base;
  a.w = b1->w;
  a.inv_ord = 1.0 / ((double)orbit_order(b1) * (double)orbit_order(b2));
  return EDK_OK;
}

template <typename VT, typename ST>
static void launch_op(const DoubleArgs &a, const VT *x, VT *y, int64_t nb, int accumulate, hipStream_t st) {
  const unsigned grid = (unsigned)((a.b1.D + 255) / 256);
  for (int64_t c = 0; c < nb;) {
    const int64_t left = nb - c;
    if (left >= 4) {
      k_double_op<VT, ST, 4><<<grid, 256, 0, st>>>(a, x, y, nb, c, accumulate);
      c += 4;
    } else if (left >= 2) {
      k_double_op<VT, ST, 2><<<grid, 256, 0, st>>>(a, x, y, nb, c, accumulate);
      c += 2;
    } else {
      k_double_op<VT, ST, 1><<<grid, 256, 0, st>>>(a, x, y, nb, c, accumulate);
      c += 1;
    }
  }
}

template <typename VT, typename ST>
static void launch_map(const DoubleArgs &a, const VT *x, VT *y, int64_t nb, hipStream_t st) {
  const unsigned grid = (unsigned)((a.b1.D + 255) / 256);
  for (int64_t c = 0; c < nb;) {
    const int64_t left = nb - c;
    if (left >= 4) {
      k_double_map<VT, ST, 4><<<grid, 256, 0, st>>>(a, x, y, nb, c);
      c += 4;
    } else if (left >= 2) {
      k_double_map<VT, ST, 2><<<grid, 256, 0, st>>>(a, x, y, nb, c);
      c += 2;
    } else {
      k_double_map<VT, ST, 1><<<grid, 256, 0, st>>>(a, x, y, nb, c);
      c += 1;
    }
  }
}

// Operator(M, I, DoubleBasis(B1, B2)) (LinearMap.jl:88-126, Operator.jl:24-78)
extern "C" int edk_double_operator_create(edk_ctx *ctx, edk_basis *b1, edk_basis *b2, int64_t nterms,
                                          const int64_t *site_off, const int64_t *sites, const int64_t *colptr_off,
                                          const int64_t *colptr, const int64_t *nz_off, const int64_t *rowval,
                                          const double *nzval, edk_dop **out) {
  if (!ctx || !out) return edk_fail(EDK_ERR_ARG, "edk_double_operator_create: null argument");
  DoubleArgs a;
  if (int rc = double_setup(b1, b2, a)) return rc;
  edk_op *op = nullptr;
  if (int rc = edk_operator_create(ctx, b1, nterms, site_off, sites, colptr_off, colptr, nz_off, rowval, nzval, &op))
    return rc;
  edk_dop *d = new edk_dop();
  d->op = op;
  d->b2 = b2;
  *out = d;
  return EDK_OK;
}

// mul!(y, H, x) / H * x for H on a DoubleBasis: y (size(B1) x nb) = (accumulate ? y : 0) + H x,
// x (size(B2) x nb), device vectors, row-interleaved columns
extern "C" int edk_double_matvec(edk_dop *d, int dtype, void *y, const void *x, int64_t nb, int accumulate) {
  if (!d || (!x && d->b2->D > 0) || (!y && d->op->basis->D > 0) || nb < 1)
    return edk_fail(EDK_ERR_ARG, "edk_double_matvec: bad argument");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "edk_double_matvec: bad dtype");
  DoubleArgs a;
  if (int rc = double_setup(d->op->basis, d->b2, a)) return rc;
  a.T = d->op->terms;
  edk_ctx *ctx = d->op->ctx;
  EDK_HIP(hipSetDevice(ctx->device));
  if (a.b1.D > 0 && a.b2.D == 0) {
    if (!accumulate) zero_out(dtype, y, a.b1.D * nb, ctx->stream);
    EDK_HIP(hipGetLastError());
  } else if (a.b1.D > 0) {
    if (dtype == EDK_C64) launch_op<double2, double>(a, (const double2 *)x, (double2 *)y, nb, accumulate, ctx->stream);
    else launch_op<float2, float>(a, (const float2 *)x, (float2 *)y, nb, accumulate, ctx->stream);
    EDK_HIP(hipGetLastError());
  }
  EDK_HIP(hipStreamSynchronize(ctx->stream));
  return EDK_OK;
}

extern "C" int edk_double_operator_destroy(edk_dop *d) {
  if (!d) return EDK_OK;
  int rc = edk_operator_destroy(d->op);
  delete d;
  return rc;
}

// (B::DoubleBasis)(v) (LinearMap.jl:144-172): out (size(B1) x nb) from v (size(B2) x nb), device
extern "C" int edk_double_map(edk_basis *b1, edk_basis *b2, int dtype, const void *v, void *out, int64_t nb) {
  DoubleArgs a;
  if (int rc = double_setup(b1, b2, a)) return rc;
  if ((!v && b2->D > 0) || (!out && b1->D > 0) || nb < 1) return edk_fail(EDK_ERR_ARG, "edk_double_map: bad argument");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "edk_double_map: bad dtype");
  edk_ctx *ctx = b1->ctx;
  EDK_HIP(hipSetDevice(ctx->device));
  if (a.b1.D > 0 && a.b2.D == 0) {
    zero_out(dtype, out, a.b1.D * nb, ctx->stream);
    EDK_HIP(hipGetLastError());
  } else if (a.b1.D > 0) {
    if (dtype == EDK_C64) launch_map<double2, double>(a, (const double2 *)v, (double2 *)out, nb, ctx->stream);
    else launch_map<float2, float>(a, (const float2 *)v, (float2 *)out, nb, ctx->stream);
    EDK_HIP(hipGetLastError());
  }
  EDK_HIP(hipStreamSynchronize(ctx->stream));
  return EDK_OK;
}
