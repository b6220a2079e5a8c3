// edk_krylov.hip -- device-resident Lanczos ground state and Krylov
// exp(-iHt)psi (KrylovEvolutionCache, TimeEvolution.jl:84-649).
//
// Vectors (the Krylov basis V, the work vector w, the anchor state) stay in
// HBM in the caller's precision (ComplexF64, or ComplexF32 for the precision
// form); every inner product, norm, reorthogonalisation coefficient and all
// tridiagonal/reduced-space arithmetic is FP64.  Reductions are two-stage
// with a grid that depends only on the vector length, and the second stage
// sums the block partials in a fixed order, so a result is bitwise the same
// whatever else runs on the device.
//
// Reorthogonalisation is classical Gram-Schmidt applied twice (CGS2) as a
// pair of GEMV sweeps over V; the reference runs the same two sweeps in
// modified-Gram-Schmidt order (TimeEvolution.jl:275-283).  Both give a basis
// orthogonal to working precision; values differ at rounding level.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "edk_krylov.h"

#define RED_TPB 256

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <typename VT>
__device__ __forceinline__ double2 ld(const VT *p, int64_t i) {
  VT v = p[i];
  return make_double2((double)v.x, (double)v.y);
}
template <typename VT>
__device__ __forceinline__ void st(VT *p, int64_t i, double2 v) {
  VT o;
  o.x = (decltype(o.x))v.x;
  o.y = (decltype(o.y))v.y;
  p[i] = o;
}

__device__ __forceinline__ double2 block_sum2(double2 v) {
  __shared__ double2 sm[RED_TPB / 64];
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_down(v.x, o, 64);
    v.y += __shfl_down(v.y, o, 64);
  }
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sm[wid] = v;
  __syncthreads();
  double2 r = make_double2(0, 0);
  if (threadIdx.x == 0)
    for (int q = 0; q < RED_TPB / 64; q++) {
      r.x += sm[q].x;
      r.y += sm[q].y;
    }
  return r;
}

// partial[b] = sum over block b's rows of conj(a) * b  (or |a|^2 when b == null)
template <typename VT>
__global__ void __launch_bounds__(RED_TPB) k_dot_part(const VT *a, const VT *b, int64_t n, double2 *part) {
  double2 acc = make_double2(0, 0);
  for (int64_t i = (int64_t)blockIdx.x * RED_TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_TPB) {
    double2 x = ld(a, i);
    if (b) {
      double2 y = ld(b, i);
      acc.x += x.x * y.x + x.y * y.y;
      acc.y += x.x * y.y - x.y * y.x;
    } else {
      acc.x += x.x * x.x + x.y * x.y;
    }
  }
  double2 r = block_sum2(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// out[c] = sum_b part[c * nb + b], fixed order; one block per c
__global__ void __launch_bounds__(RED_TPB) k_reduce_part(const double2 *part, int nb, double2 *out) {
  double2 acc = make_double2(0, 0);
  const double2 *p = part + (int64_t)blockIdx.x * nb;
  for (int b = threadIdx.x; b < nb; b += RED_TPB) {
    acc.x += p[b].x;
    acc.y += p[b].y;
  }
  double2 r = block_sum2(acc);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

// the value a store of v into a VT vector keeps
template <typename VT>
__device__ __forceinline__ double2 rnd(double2 v) {
  VT o;
  o.x = (decltype(o.x))v.x;
  o.y = (decltype(o.y))v.y;
  return make_double2((double)o.x, (double)o.y);
}

// Fused axpy + reduction, one pass: w += alpha * x, then partial[b] = block
// b's share of conj(y) . w (|w|^2 when y == null) over the stored w.  Same
// rows per block and the same reduction tree as k_dot_part.
template <typename VT>
__global__ void __launch_bounds__(RED_TPB) k_axpy_dot_part(VT *w, const VT *x, double2 alpha, const double2 *dalpha,
                                                           double sign, const VT *y, int64_t n, double2 *part) {
  // dalpha != null: alpha = sign * Re(dalpha[0]), a coefficient left on the device
  if (dalpha) alpha = make_double2(sign * dalpha[0].x, 0.0);
  double2 acc = make_double2(0, 0);
  for (int64_t i = (int64_t)blockIdx.x * RED_TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_TPB) {
    double2 a = ld(w, i), b = ld(x, i);
    a.x += alpha.x * b.x - alpha.y * b.y;
    a.y += alpha.x * b.y + alpha.y * b.x;
    st(w, i, a);
    a = rnd<VT>(a);
    if (y) {
      const double2 c = ld(y, i);
      acc.x += c.x * a.x + c.y * a.y;
      acc.y += c.x * a.y - c.y * a.x;
    } else {
      acc.x += a.x * a.x + a.y * a.y;
    }
  }
  double2 r = block_sum2(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// y += sign * Re(dalpha[0]) * x
template <typename VT>
__global__ void k_axpy_re_dev(VT *y, const VT *x, const double2 *dalpha, double sign, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = sign * dalpha[0].x;
  double2 u = ld(y, i), b = ld(x, i);
  u.x += a * b.x - 0.0 * b.y;
  u.y += a * b.y + 0.0 * b.x;
  st(y, i, u);
}

// y += alpha * x
template <typename VT>
__global__ void k_axpy(VT *y, const VT *x, double2 alpha, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 a = ld(y, i), b = ld(x, i);
  a.x += alpha.x * b.x - alpha.y * b.y;
  a.y += alpha.x * b.y + alpha.y * b.x;
  st(y, i, a);
}

// y = x * s  (real s)
template <typename VT, typename VS>
__global__ void k_scale(VT *y, const VS *x, double s, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 b = ld(x, i);
  st(y, i, make_double2(b.x * s, b.y * s));
}

// y = x * Re(sp[0])  (a scale left on the device)
template <typename VT>
__global__ void k_scale_dev(VT *y, const VT *x, const double2 *sp, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = sp[0].x;
  double2 b = ld(x, i);
  st(y, i, make_double2(b.x * s, b.y * s));
}

// Device-driven Lanczos step j (KrylovWork::d_lz layout): the norm partials
// of w after the first Gram-Schmidt sweep are reduced as k_reduce_part does,
// then thread 0 takes the host loop's decisions on the same doubles: the
// second-sweep test ||w'||^2 < (||w'||^2 + ||h||^2) / 2 over the sweep's
// coefficients h = s[0..j), beta_j = sqrt(||w'||^2), the breakdown rule and
// 1/beta_j.  The first step that wants a second sweep (reason 1) or breaks
// down (reason 2) is recorded; the host takes over from there.
#define LZ_M EDK_KRYLOV_MAXM
__global__ void __launch_bounds__(RED_TPB) k_lz_step(const double2 *part, int nb, const double2 *s, int j,
                                                     double2 *lz) {
  double2 acc = make_double2(0, 0);
  for (int b = threadIdx.x; b < nb; b += RED_TPB) {
    acc.x += part[b].x;
    acc.y += part[b].y;
  }
  const double2 r = block_sum2(acc);
  if (threadIdx.x == 0) {
    const double w2 = r.x;
    double hh = 0.0;
    for (int c = 0; c < j; c++) hh += s[c].x * s[c].x + s[c].y * s[c].y;
    const bool second = w2 < 0.5 * (w2 + hh);
    const double bj = sqrt(w2), aj = lz[j - 1].x;
    lz[LZ_M + j - 1] = make_double2(bj, second ? 1.0 : 0.0);
    lz[2 * LZ_M] = make_double2(1.0 / bj, 0.0);
    const double sq = 1.4901161193847656e-08;  // sqrt(eps(Float64))
    if (lz[2 * LZ_M + 1].x == 0.0) {
      if (second) lz[2 * LZ_M + 1] = make_double2((double)j, 1.0);
      else if (bj <= sq * fmax(1.0, fabs(aj))) lz[2 * LZ_M + 1] = make_double2((double)j, 2.0);
    }
  }
}

// For small vectors, the device-driven steps fold the one-block reductions
// into the kernels that consume them: every block sums the partials in
// k_reduce_part's order and tree (so each holds the same doubles), block 0
// publishes the result.  Saves a launch per reduction where the blocks'
// repeated reads of the partials cost less than a launch.
//
// w -= Re(alpha_j) v_j with alpha_j the sum of the partials, written to *out
template <typename VT>
__global__ void __launch_bounds__(RED_TPB) k_axpy_re_red(VT *y, const VT *x, const double2 *part, int nb,
                                                         double2 *out, double sign, int64_t n) {
  __shared__ double2 sa;
  double2 acc = make_double2(0, 0);
  for (int b = threadIdx.x; b < nb; b += RED_TPB) {
    acc.x += part[b].x;
    acc.y += part[b].y;
  }
  const double2 r = block_sum2(acc);
  if (threadIdx.x == 0) {
    sa = r;
    if (blockIdx.x == 0) out[0] = r;
  }
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * RED_TPB + threadIdx.x;
  if (i >= n) return;
  const double a = sign * sa.x;
  double2 u = ld(y, i), b = ld(x, i);
  u.x += a * b.x - 0.0 * b.y;
  u.y += a * b.y + 0.0 * b.x;
  st(y, i, u);
}

// k_lz_step and k_scale_dev in one launch: every block takes beta_j from the
// norm partials as k_lz_step does, block 0 records the step's scalars and
// decisions, and (scale != 0) the blocks write v_{j+1} = w / beta_j
template <typename VT>
__global__ void __launch_bounds__(RED_TPB) k_lz_step_scale(const double2 *part, int nb, const double2 *s, int j,
                                                           double2 *lz, VT *vn, const VT *w, int64_t n,
                                                           int scale) {
  __shared__ double sinv;
  double2 acc = make_double2(0, 0);
  for (int b = threadIdx.x; b < nb; b += RED_TPB) {
    acc.x += part[b].x;
    acc.y += part[b].y;
  }
  const double2 r = block_sum2(acc);
  if (threadIdx.x == 0) {
    const double w2 = r.x;
    const double bj = sqrt(w2);
    sinv = 1.0 / bj;
    if (blockIdx.x == 0) {
      double hh = 0.0;
      for (int c = 0; c < j; c++) hh += s[c].x * s[c].x + s[c].y * s[c].y;
      const bool second = w2 < 0.5 * (w2 + hh);
      const double aj = lz[j - 1].x;
      lz[LZ_M + j - 1] = make_double2(bj, second ? 1.0 : 0.0);
      lz[2 * LZ_M] = make_double2(1.0 / bj, 0.0);
      const double sq = 1.4901161193847656e-08;  // sqrt(eps(Float64))
      if (lz[2 * LZ_M + 1].x == 0.0) {
        if (second) lz[2 * LZ_M + 1] = make_double2((double)j, 1.0);
        else if (bj <= sq * fmax(1.0, fabs(aj))) lz[2 * LZ_M + 1] = make_double2((double)j, 2.0);
      }
    }
  }
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * RED_TPB + threadIdx.x;
  if (!scale || i >= n) return;
  const double sc = sinv;
  double2 b = ld(w, i);
  st(vn, i, make_double2(b.x * sc, b.y * sc));
}

// start of a device-driven extension from step m0: no step settled yet, and
// beta_{m0} (from the host's array) where step m0 + 1 reads it
__global__ void k_lz_start(double2 *lz, int m0, double beta_m0) {
  lz[2 * LZ_M + 1] = make_double2(0.0, 0.0);
  if (m0 >= 1) lz[LZ_M + m0 - 1] = make_double2(beta_m0, 0.0);
}

// part[c * nblk + b] = block b's share of conj(V[:, c]) . w,  c < j.
// Columns go in groups of GEMV_CG: each thread loads w_i once per group and
// keeps one accumulator per column, so w is read j / GEMV_CG times instead of
// j times; the group's block sums are reduced together (wave shuffles, then
// one LDS pass over the waves), in a fixed order.
#define GEMV_CG 8
template <typename VT>
__global__ void __launch_bounds__(RED_TPB) k_gemv_t(const VT *V, int64_t ld_v, int j, const VT *w, int64_t n,
                                                    double2 *part) {
  __shared__ double2 sm[RED_TPB / 64][GEMV_CG];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int c0 = 0; c0 < j; c0 += GEMV_CG) {
    const int cg = min(GEMV_CG, j - c0);
    double2 acc[GEMV_CG];
#pragma unroll
    for (int q = 0; q < GEMV_CG; q++) acc[q] = make_double2(0, 0);
    for (int64_t i = (int64_t)blockIdx.x * RED_TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_TPB) {
      const double2 y = ld(w, i);
#pragma unroll
      for (int q = 0; q < GEMV_CG; q++) {
        if (q < cg) {
          const double2 x = ld(V + (int64_t)(c0 + q) * ld_v, i);
          acc[q].x += x.x * y.x + x.y * y.y;
          acc[q].y += x.x * y.y - x.y * y.x;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < GEMV_CG; q++) {
      for (int o = 32; o > 0; o >>= 1) {
        acc[q].x += __shfl_down(acc[q].x, o, 64);
        acc[q].y += __shfl_down(acc[q].y, o, 64);
      }
    }
    __syncthreads();
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < GEMV_CG; q++) sm[wid][q] = acc[q];
    }
    __syncthreads();
    if (threadIdx.x < cg) {
      double2 r = make_double2(0, 0);
      for (int u = 0; u < RED_TPB / 64; u++) {
        r.x += sm[u][threadIdx.x].x;
        r.y += sm[u][threadIdx.x].y;
      }
      part[(int64_t)(c0 + threadIdx.x) * gridDim.x + blockIdx.x] = r;
    }
  }
}

// k_gemv_t for a ComplexF64 basis with w read once: each block walks row
// tiles of GEMV_TR rows (GEMV_RPT per lane, held in registers; 8 measured
// faster than 4, and both faster than re-reading w per column group), and for each
// tile runs every column group against them, adding the group's block sum
// into its own partial slots (the first tile stores).  The slots of a block
// are written only by that block, in tile order: no atomics, a fixed order.
template <int GEMV_RPT>
__global__ void __launch_bounds__(RED_TPB) k_gemv_t_tile(const double2 *V, int64_t ld_v, int j, const double2 *w,
                                                         int64_t n, double2 *part) {
  constexpr int64_t GEMV_TR = RED_TPB * GEMV_RPT;
  __shared__ double2 sm[RED_TPB / 64][GEMV_CG];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t ntile = (n + GEMV_TR - 1) / GEMV_TR;
  bool first = true;
  for (int64_t t = blockIdx.x; t < ntile; t += gridDim.x, first = false) {
    const int64_t r0 = t * GEMV_TR + threadIdx.x;
    double2 y[GEMV_RPT];
#pragma unroll
    for (int k = 0; k < GEMV_RPT; k++) {
      const int64_t i = r0 + (int64_t)k * RED_TPB;
      y[k] = i < n ? w[i] : make_double2(0, 0);
    }
    for (int c0 = 0; c0 < j; c0 += GEMV_CG) {
      const int cg = min(GEMV_CG, j - c0);
      double2 acc[GEMV_CG];
#pragma unroll
      for (int q = 0; q < GEMV_CG; q++) {
        acc[q] = make_double2(0, 0);
        if (q < cg) {
          const double2 *col = V + (int64_t)(c0 + q) * ld_v;
#pragma unroll
          for (int k = 0; k < GEMV_RPT; k++) {
            const int64_t i = r0 + (int64_t)k * RED_TPB;
            const double2 x = i < n ? col[i] : make_double2(0, 0);
            acc[q].x += x.x * y[k].x + x.y * y[k].y;
            acc[q].y += x.x * y[k].y - x.y * y[k].x;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < GEMV_CG; q++) {
        for (int o = 32; o > 0; o >>= 1) {
          acc[q].x += __shfl_down(acc[q].x, o, 64);
          acc[q].y += __shfl_down(acc[q].y, o, 64);
        }
      }
      __syncthreads();
      if (lane == 0) {
#pragma unroll
        for (int q = 0; q < GEMV_CG; q++) sm[wid][q] = acc[q];
      }
      __syncthreads();
      if (threadIdx.x < cg) {
        double2 r = make_double2(0, 0);
        for (int u = 0; u < RED_TPB / 64; u++) {
          r.x += sm[u][threadIdx.x].x;
          r.y += sm[u][threadIdx.x].y;
        }
        double2 *slot = part + (int64_t)(c0 + threadIdx.x) * gridDim.x + blockIdx.x;
        if (!first) {
          const double2 prev = *slot;
          r.x = prev.x + r.x;
          r.y = prev.y + r.y;
        }
        *slot = r;
      }
    }
  }
  // a block without a tile (more blocks than tiles) leaves zeros
  if (blockIdx.x >= ntile && threadIdx.x < RED_TPB)
    for (int c = threadIdx.x; c < j; c += RED_TPB) part[(int64_t)c * gridDim.x + blockIdx.x] = make_double2(0, 0);
}

// k_gemv_t for a ComplexF32 basis: two rows per lane with 16-byte loads (one
// 8-byte load per lane per column leaves the kernel far below the copy rate),
// FP64 accumulation.  Needs an even leading dimension and 16-byte aligned
// V and w; an odd last row is added by lane 0 of block 0.
__global__ void __launch_bounds__(RED_TPB) k_gemv_t_c32x2(const float2 *V, int64_t ld_v, int j, const float2 *w,
                                                          int64_t n, double2 *part) {
  __shared__ double2 sm[RED_TPB / 64][GEMV_CG];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t n2 = n >> 1;
  const bool tail = (n & 1) && blockIdx.x == 0 && threadIdx.x == 0;
  for (int c0 = 0; c0 < j; c0 += GEMV_CG) {
    const int cg = min(GEMV_CG, j - c0);
    double2 acc[GEMV_CG];
#pragma unroll
    for (int q = 0; q < GEMV_CG; q++) acc[q] = make_double2(0, 0);
    for (int64_t p = (int64_t)blockIdx.x * RED_TPB + threadIdx.x; p < n2; p += (int64_t)gridDim.x * RED_TPB) {
      const float4 y = reinterpret_cast<const float4 *>(w)[p];
#pragma unroll
      for (int q = 0; q < GEMV_CG; q++) {
        if (q < cg) {
          const float4 x = reinterpret_cast<const float4 *>(V + (int64_t)(c0 + q) * ld_v)[p];
          acc[q].x += (double)x.x * y.x + (double)x.y * y.y;
          acc[q].y += (double)x.x * y.y - (double)x.y * y.x;
          acc[q].x += (double)x.z * y.z + (double)x.w * y.w;
          acc[q].y += (double)x.z * y.w - (double)x.w * y.z;
        }
      }
    }
    if (tail) {
      const float2 y = w[n - 1];
#pragma unroll
      for (int q = 0; q < GEMV_CG; q++) {
        if (q < cg) {
          const float2 x = V[(int64_t)(c0 + q) * ld_v + n - 1];
          acc[q].x += (double)x.x * y.x + (double)x.y * y.y;
          acc[q].y += (double)x.x * y.y - (double)x.y * y.x;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < GEMV_CG; q++) {
      for (int o = 32; o > 0; o >>= 1) {
        acc[q].x += __shfl_down(acc[q].x, o, 64);
        acc[q].y += __shfl_down(acc[q].y, o, 64);
      }
    }
    __syncthreads();
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < GEMV_CG; q++) sm[wid][q] = acc[q];
    }
    __syncthreads();
    if (threadIdx.x < cg) {
      double2 r = make_double2(0, 0);
      for (int u = 0; u < RED_TPB / 64; u++) {
        r.x += sm[u][threadIdx.x].x;
        r.y += sm[u][threadIdx.x].y;
      }
      part[(int64_t)(c0 + threadIdx.x) * gridDim.x + blockIdx.x] = r;
    }
  }
}

// out = scale_in * out_in - sum_c V[:, c] s[c]   (mode 0: w -= V s)
// out = scale * sum_c V[:, c] s[c]               (mode 1: reconstruct)
template <typename VT, typename VO>
__global__ void __launch_bounds__(RED_TPB) k_gemv_n(const VT *V, int64_t ld_v, int j, const double2 *s, VO *out,
                                                    int64_t n, int mode, double scale) {
  __shared__ double2 sc[EDK_KRYLOV_MAXM + 1];
  for (int c = threadIdx.x; c < j; c += RED_TPB) sc[c] = s[c];
  __syncthreads();
  int64_t i = (int64_t)blockIdx.x * RED_TPB + threadIdx.x;
  if (i >= n) return;
  double2 acc = make_double2(0, 0);
  for (int c = 0; c < j; c++) {
    double2 v = ld(V + (int64_t)c * ld_v, i);
    acc.x += v.x * sc[c].x - v.y * sc[c].y;
    acc.y += v.x * sc[c].y + v.y * sc[c].x;
  }
  if (mode == 0) {
    double2 o = ld(out, i);
    st(out, i, make_double2(o.x - acc.x, o.y - acc.y));
  } else {
    st(out, i, make_double2(acc.x * scale, acc.y * scale));
  }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
static int red_blocks(int64_t n) {
  int64_t b = (n + RED_TPB - 1) / RED_TPB;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, 1024));
}

int KrylovWork::init(edk_ctx *c, int64_t n_, int dtype_, int maxm, int slot) {
  ctx = c;
  n = n_;
  dtype = dtype_;
  nblk = red_blocks(n);
  if (maxm + 2 > EDK_PINNED_SLOT) return edk_fail(EDK_ERR_UNSUPPORTED, "Krylov dimension above 1024");
  size_t pb = (size_t)nblk * (maxm + 2) * sizeof(double2);
  if (edk_dalloc(ctx, (void **)&d_part, pb) != hipSuccess) return edk_fail(EDK_ERR_OOM, "krylov partials");
  // coefficients in d_s[0..maxm+1], a norm in d_s[NORM_SLOT]
  static_assert(NORM_SLOT < EDK_PINNED_SLOT, "norm slot must fit in the pinned buffer");
  static_assert(ALPHA_SLOT < EDK_PINNED_SLOT, "coefficient slot must fit in the pinned buffer");
  if (edk_dalloc(ctx, (void **)&d_s, (NORM_SLOT + 1) * sizeof(double2)) != hipSuccess)
    return edk_fail(EDK_ERR_OOM, "krylov coefficients");
  if (edk_dalloc(ctx, (void **)&d_lz, (2 * EDK_KRYLOV_MAXM + 2) * sizeof(double2)) != hipSuccess)
    return edk_fail(EDK_ERR_OOM, "krylov step scalars");
  h_s = (double2 *)ctx->pinned + (size_t)slot * EDK_PINNED_SLOT;
  return EDK_OK;
}

void KrylovWork::release() {
  if (ctx) {
    edk_dfree(ctx, d_part);
    edk_dfree(ctx, d_s);
    edk_dfree(ctx, d_lz);
  }
  d_part = nullptr;
  d_s = nullptr;
  d_lz = nullptr;
  h_s = nullptr;
}

size_t KrylovWork::esize() const { return dtype == EDK_C64 ? 16 : 8; }

#define DISPATCH(VTNAME, ...)                      \
  do {                                             \
    if (dtype == EDK_C64) {                        \
      using VTNAME = double2;                      \
      __VA_ARGS__;                                 \
    } else {                                       \
      using VTNAME = float2;                       \
      __VA_ARGS__;                                 \
    }                                              \
  } while (0)

int KrylovWork::dot(const void *a, const void *b, double2 *out) {
  hipStream_t s = ctx->stream;
  DISPATCH(VT, k_dot_part<VT><<<nblk, RED_TPB, 0, s>>>((const VT *)a, (const VT *)b, n, d_part));
  k_reduce_part<<<1, RED_TPB, 0, s>>>(d_part, nblk, d_s);
  EDK_HIP(hipGetLastError());
  EDK_HIP(hipMemcpyAsync(h_s, d_s, sizeof(double2), hipMemcpyDeviceToHost, s));
  EDK_HIP(hipStreamSynchronize(s));
  *out = h_s[0];
  return EDK_OK;
}

int KrylovWork::axpy_dot(void *w, const void *x, double2 alpha, const void *y, double2 *out) {
  hipStream_t s = ctx->stream;
  DISPATCH(VT, k_axpy_dot_part<VT><<<nblk, RED_TPB, 0, s>>>((VT *)w, (const VT *)x, alpha, nullptr, 0.0,
                                                            (const VT *)y, n, d_part));
  k_reduce_part<<<1, RED_TPB, 0, s>>>(d_part, nblk, d_s);
  EDK_HIP(hipGetLastError());
  EDK_HIP(hipMemcpyAsync(h_s, d_s, sizeof(double2), hipMemcpyDeviceToHost, s));
  EDK_HIP(hipStreamSynchronize(s));
  *out = h_s[0];
  return EDK_OK;
}

// The Lanczos coefficient stays on the device: conj(y) . (w + alpha x) (or
// conj(y) . w when x == null) reduced into d_s[0], and copied to
// h_s[ALPHA_SLOT] in stream order (read it after the next synchronisation).
int KrylovWork::coef_async(void *w, const void *x, double2 alpha, const void *y) {
  hipStream_t s = ctx->stream;
  if (x) {
    DISPATCH(VT, k_axpy_dot_part<VT><<<nblk, RED_TPB, 0, s>>>((VT *)w, (const VT *)x, alpha, nullptr, 0.0,
                                                              (const VT *)y, n, d_part));
  } else {
    DISPATCH(VT, k_dot_part<VT><<<nblk, RED_TPB, 0, s>>>((const VT *)y, (const VT *)w, n, d_part));
  }
  k_reduce_part<<<1, RED_TPB, 0, s>>>(d_part, nblk, d_s);
  EDK_HIP(hipGetLastError());
  EDK_HIP(hipMemcpyAsync(h_s + ALPHA_SLOT, d_s, sizeof(double2), hipMemcpyDeviceToHost, s));
  return EDK_OK;
}

// w -= Re(d_s[0]) x, then (norm_out != null) |w|^2, synchronised
int KrylovWork::sub_coef(void *w, const void *x, double *norm2_out) {
  hipStream_t s = ctx->stream;
  if (norm2_out) {
    DISPATCH(VT, k_axpy_dot_part<VT><<<nblk, RED_TPB, 0, s>>>((VT *)w, (const VT *)x, make_double2(0, 0), d_s, -1.0,
                                                              (const VT *)nullptr, n, d_part));
    k_reduce_part<<<1, RED_TPB, 0, s>>>(d_part, nblk, d_s + NORM_SLOT);
    EDK_HIP(hipGetLastError());
    EDK_HIP(hipMemcpyAsync(h_s + NORM_SLOT, d_s + NORM_SLOT, sizeof(double2), hipMemcpyDeviceToHost, s));
    EDK_HIP(hipStreamSynchronize(s));
    *norm2_out = h_s[NORM_SLOT].x;
  } else {
    unsigned g = (unsigned)((n + 255) / 256);
    DISPATCH(VT, k_axpy_re_dev<VT><<<g, 256, 0, s>>>((VT *)w, (const VT *)x, d_s, -1.0, n));
    EDK_HIP(hipGetLastError());
  }
  return EDK_OK;
}

int KrylovWork::norm(const void *a, double *out) {
  double2 r;
  if (int rc = dot(a, nullptr, &r)) return rc;
  *out = std::sqrt(r.x);
  return EDK_OK;
}

int KrylovWork::axpy(void *y, const void *x, double2 alpha) {
  unsigned g = (unsigned)((n + 255) / 256);
  DISPATCH(VT, k_axpy<VT><<<g, 256, 0, ctx->stream>>>((VT *)y, (const VT *)x, alpha, n));
  EDK_HIP(hipGetLastError());
  return EDK_OK;
}

int KrylovWork::scale(void *y, const void *x, double s) {
  unsigned g = (unsigned)((n + 255) / 256);
  DISPATCH(VT, k_scale<VT, VT><<<g, 256, 0, ctx->stream>>>((VT *)y, (const VT *)x, s, n));
  EDK_HIP(hipGetLastError());
  return EDK_OK;
}

// One classical Gram-Schmidt sweep: h = V[:, :j]^H w, w -= V[:, :j] h.
int KrylovWork::cgs_sweep(const void *V, int64_t ldv, int j, void *w) {
  hipStream_t s = ctx->stream;
  if (dtype == EDK_C32 && ldv % 2 == 0 && ((uintptr_t)V & 15) == 0 && ((uintptr_t)w & 15) == 0)
    k_gemv_t_c32x2<<<nblk, RED_TPB, 0, s>>>((const float2 *)V, ldv, j, (const float2 *)w, n, d_part);
  else if (dtype == EDK_C64 && n >= (int64_t)RED_TPB * 8 * nblk)
    // every block has at least one tile (about 2.1 M rows and up); below
    // that most blocks would idle (L = 24 evolution: 2120 -> 2000 outputs/s)
    k_gemv_t_tile<8><<<nblk, RED_TPB, 0, s>>>((const double2 *)V, ldv, j, (const double2 *)w, n, d_part);
  else
    DISPATCH(VT, k_gemv_t<VT><<<nblk, RED_TPB, 0, s>>>((const VT *)V, ldv, j, (const VT *)w, n, d_part));
  k_reduce_part<<<j, RED_TPB, 0, s>>>(d_part, nblk, d_s);
  unsigned g = (unsigned)((n + RED_TPB - 1) / RED_TPB);
  DISPATCH(VT, k_gemv_n<VT, VT><<<g, RED_TPB, 0, s>>>((const VT *)V, ldv, j, d_s, (VT *)w, n, 0, 0.0));
  EDK_HIP(hipGetLastError());
  return EDK_OK;
}

// Reorthogonalise w against V[:, :j] and return ||w||: a first CGS sweep,
// then a second one only when the first removed a large part of w (the
// criterion of Daniel, Gragg, Kaufman and Stewart: ||w'|| < ||w|| / sqrt(2),
// with ||w||^2 = ||w'||^2 + ||h||^2 since w' is orthogonal to V h).  In a
// Lanczos step w is already nearly orthogonal to V, so the second sweep --
// and its two passes over V -- is rarely needed; when it is, two sweeps
// leave w orthogonal to working precision ("twice is enough").
int KrylovWork::reorth_norm(const void *V, int64_t ldv, int j, void *w, double *wnorm, int *sweeps, double *hmax) {
  hipStream_t s = ctx->stream;
  if (int rc = cgs_sweep(V, ldv, j, w)) return rc;
  // h back to the host while the norm of w' is reduced (one synchronisation)
  EDK_HIP(hipMemcpyAsync(h_s, d_s, (size_t)j * sizeof(double2), hipMemcpyDeviceToHost, s));
  DISPATCH(VT, k_dot_part<VT><<<nblk, RED_TPB, 0, s>>>((const VT *)w, (const VT *)nullptr, n, d_part));
  k_reduce_part<<<1, RED_TPB, 0, s>>>(d_part, nblk, d_s + NORM_SLOT);
  EDK_HIP(hipGetLastError());
  EDK_HIP(hipMemcpyAsync(h_s + NORM_SLOT, d_s + NORM_SLOT, sizeof(double2), hipMemcpyDeviceToHost, s));
  EDK_HIP(hipStreamSynchronize(s));
  double hh = 0.0, hm = 0.0;
  for (int c = 0; c < j; c++) {
    const double a = h_s[c].x * h_s[c].x + h_s[c].y * h_s[c].y;
    hh += a;
    hm = std::max(hm, a);
  }
  double w2 = h_s[NORM_SLOT].x;
  if (hmax) *hmax = std::sqrt(hm / std::max(w2 + hh, 1e-300));
  int nsw = 1;
  if (w2 < 0.5 * (w2 + hh)) {
    if (int rc = cgs_sweep(V, ldv, j, w)) return rc;
    double nrm;
    if (int rc = norm(w, &nrm)) return rc;
    w2 = nrm * nrm;
    nsw = 2;
  }
  *wnorm = std::sqrt(w2);
  if (sweeps) *sweeps = nsw;
  return EDK_OK;
}

// out = scale * V[:, :j] c   (c: host complex coefficients)
int KrylovWork::combine(const void *V, int64_t ldv, int j, const double2 *c, double scale, void *out, int out_dtype) {
  hipStream_t s = ctx->stream;
  memcpy(h_s, c, j * sizeof(double2));
  EDK_HIP(hipMemcpyAsync(d_s, h_s, j * sizeof(double2), hipMemcpyHostToDevice, s));
  unsigned g = (unsigned)((n + RED_TPB - 1) / RED_TPB);
  if (out_dtype == EDK_C64) {
    DISPATCH(VT, k_gemv_n<VT, double2><<<g, RED_TPB, 0, s>>>((const VT *)V, ldv, j, d_s, (double2 *)out, n, 1, scale));
  } else {
    DISPATCH(VT, k_gemv_n<VT, float2><<<g, RED_TPB, 0, s>>>((const VT *)V, ldv, j, d_s, (float2 *)out, n, 1, scale));
  }
  EDK_HIP(hipGetLastError());
  // the host buffer is reused by the next call: wait for the copy
  EDK_HIP(hipStreamSynchronize(s));
  return EDK_OK;
}

// ---------------------------------------------------------------------------
// symmetric tridiagonal eigensolver (implicit QL with Wilkinson shift, the
// textbook method of Golub & Van Loan 8.3), FP64
// ---------------------------------------------------------------------------
// Z receives rows `rows[0..nr-1]` of the eigenvector matrix, column-major
// (Z[r + i*nr] = component rows[r] of eigenvector i); rows == null means all
// n rows (nr = n).  The rotations act on each row independently, so tracking
// a few rows costs O(n) per rotation sweep instead of O(n^2).
int edk_tridiag_eig(int n, const double *alpha, const double *beta, double *d, double *Z, int nr,
                    const int *rows) {
  if (!rows) nr = n;
  std::vector<double> e(n + 1, 0.0);
  for (int i = 0; i < n; i++) d[i] = alpha[i];
  for (int i = 0; i + 1 < n; i++) e[i] = beta[i];
  for (int i = 0; i < nr * n; i++) Z[i] = 0.0;
  for (int r = 0; r < nr; r++) Z[r + (size_t)(rows ? rows[r] : r) * nr] = 1.0;
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < n; l++) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; m++) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m != l) {
        if (iter++ == 100) return edk_fail(EDK_ERR_NUMERIC, "tridiagonal eigensolver did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; i--) {
          double f = s * e[i], bb = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * bb;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - bb;
          for (int k = 0; k < nr; k++) {
            double zf = Z[k + (size_t)(i + 1) * nr];
            Z[k + (size_t)(i + 1) * nr] = s * Z[k + (size_t)i * nr] + c * zf;
            Z[k + (size_t)i * nr] = c * Z[k + (size_t)i * nr] - s * zf;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  for (int i = 0; i < n - 1; i++) {
    int k = i;
    for (int j = i + 1; j < n; j++)
      if (d[j] < d[k]) k = j;
    if (k != i) {
      std::swap(d[i], d[k]);
      for (int q = 0; q < nr; q++) std::swap(Z[q + (size_t)i * nr], Z[q + (size_t)k * nr]);
    }
  }
  return EDK_OK;
}

// Lowest eigenvalue of the symmetric tridiagonal (alpha, beta) and the
// magnitude of the last component of its normalised eigenvector, in O(n):
// bisection on Sturm counts (Golub & Van Loan 8.4.1) to full precision, then
// two steps of inverse iteration on T - theta I with a row-interchanging LU
// (the method of LAPACK's dstein).  The per-step convergence test of the
// Lanczos ground state needs only these two numbers; the QL solver above is
// O(n^2) per call.
static int sturm_below(int n, const double *a, const double *b, double x) {
  int c = 0;
  double d = 1.0;
  for (int k = 0; k < n; k++) {
    d = (a[k] - x) - (k ? b[k - 1] * b[k - 1] / d : 0.0);
    if (d == 0.0) d = -1e-300;
    if (d < 0) c++;
  }
  return c;
}

static void lowest_ritz(int n, const double *a, const double *b, double *theta, double *qlast) {
  double lo = INFINITY, hi = -INFINITY, tn = 0.0;
  for (int k = 0; k < n; k++) {
    const double r = (k ? std::fabs(b[k - 1]) : 0.0) + (k < n - 1 ? std::fabs(b[k]) : 0.0);
    lo = std::min(lo, a[k] - r);
    hi = std::max(hi, a[k] + r);
    tn = std::max(tn, std::fabs(a[k]) + r);
  }
  for (int it = 0; it < 256; it++) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm_below(n, a, b, mid) >= 1) hi = mid;
    else lo = mid;
  }
  const double th = hi;
  *theta = th;
  // LU with row interchanges of A = T - th I: U has diagonal d, superdiagonals du, du2
  std::vector<double> d(n), du(n, 0.0), du2(n, 0.0), dl(n, 0.0), x(n, 1.0);
  std::vector<char> piv(n, 0);
  const double pert = 2.220446049250313e-16 * std::max(tn, 1e-300);
  for (int k = 0; k < n; k++) {
    d[k] = a[k] - th;
    if (k < n - 1) du[k] = dl[k] = b[k];
  }
  for (int k = 0; k + 1 < n; k++) {
    if (std::fabs(d[k]) >= std::fabs(dl[k])) {
      if (d[k] == 0.0) d[k] = pert;
      const double l = dl[k] / d[k];
      dl[k] = l;
      d[k + 1] -= l * du[k];
    } else {
      const double l = d[k] / dl[k];
      d[k] = dl[k];
      dl[k] = l;
      const double t = du[k];
      du[k] = d[k + 1];
      d[k + 1] = t - l * d[k + 1];
      if (k + 2 < n) {
        du2[k] = du[k + 1];
        du[k + 1] = -l * du2[k];
      }
      piv[k] = 1;
    }
  }
  if (d[n - 1] == 0.0) d[n - 1] = pert;
  for (int it = 0; it < 2; it++) {
    for (int k = 0; k + 1 < n; k++) {
      if (!piv[k]) x[k + 1] -= dl[k] * x[k];
      else {
        const double t = x[k];
        x[k] = x[k + 1];
        x[k + 1] = t - dl[k] * x[k];
      }
    }
    x[n - 1] /= d[n - 1];
    if (n > 1) x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / d[n - 2];
    for (int k = n - 3; k >= 0; k--) x[k] = (x[k] - du[k] * x[k + 1] - du2[k] * x[k + 2]) / d[k];
    double mx = 0.0;
    for (int k = 0; k < n; k++) mx = std::max(mx, std::fabs(x[k]));
    for (int k = 0; k < n; k++) x[k] /= mx;
  }
  double s2 = 0.0;
  for (int k = 0; k < n; k++) s2 += x[k] * x[k];
  *qlast = std::fabs(x[n - 1]) / std::sqrt(s2);
}

// ---------------------------------------------------------------------------
// Lanczos step: the _lanczos_build! recurrence (TimeEvolution.jl:250-306)
// ---------------------------------------------------------------------------
bool PartialOrth::update(int j, const double *alpha, const double *beta, double bj) {
  const double aj = alpha[j - 1], bjm1 = j > 1 ? beta[j - 2] : 0.0;
  anorm = std::max(anorm, std::fabs(aj) + bjm1 + bj);
  const double th = 10.0 * eps * anorm;
  std::vector<double> wn(j + 1);
  double mx = 0.0;
  if (bj > 0) {
    for (int k = 1; k <= j - 1; k++) {
      double t = beta[k - 1] * wcur[k] + (alpha[k - 1] - aj) * wcur[k - 1] - bjm1 * wold[k - 1];
      if (k > 1) t += beta[k - 2] * wcur[k - 2];
      t = (t + std::copysign(th, t)) / bj;
      wn[k - 1] = t;
      mx = std::max(mx, std::fabs(t));
    }
    wn[j - 1] = th / bj;  // consecutive vectors: the local alpha/beta removal
  } else {
    mx = INFINITY;
  }
  wn[j] = 1.0;
  wold.swap(wcur);
  wcur.swap(wn);
  forced_now = force;
  return force || !(mx <= std::sqrt(eps));
}

void PartialOrth::reset(int j) {
  for (int k = 0; k < j; k++) wcur[k] = eps;
  force = !forced_now;  // a threshold crossing also reorthogonalises the next vector
  reorths++;
}

// Extends the basis V (columns 0..) from m_start to at most m_target steps.
// Returns 1 on lucky breakdown (beta set to 0), 0 otherwise; <0 on error.
// po == nullptr: full reorthogonalisation at every step (the reference's
// two-sweep Gram-Schmidt); otherwise partial reorthogonalisation.
static int lanczos_extend_host(edk_op *op, KrylovWork &K, void *V, int64_t ldv, void *w, double *alpha,
                               double *beta, int *m, int m_target, int64_t *matvecs, int ncols, PartialOrth *po,
                               double brk) {
  size_t es = K.esize();
  for (int j = *m + 1; j <= m_target; j++) {
    char *vj = (char *)V + (size_t)(j - 1) * ldv * es;
    if (int rc = edk_matvec_launch(op, K.dtype, w, vj, 1, 0, K.ctx->stream)) return -rc;
    (*matvecs)++;
    // w -= beta_{j-1} v_{j-1} and alpha_j = v_j^H w in one pass; alpha_j
    // stays on the device for the next update (one host synchronisation per
    // step, at the norm), its host copy arriving in stream order
    char *vjm1 = j > 1 ? (char *)V + (size_t)(j - 2) * ldv * es : nullptr;
    if (int rc = K.coef_async(w, vjm1, make_double2(j > 1 ? -beta[j - 2] : 0.0, 0.0), vj)) return -rc;
    double bj;
    if (!po) {
      if (int rc = K.sub_coef(w, vj, nullptr)) return -rc;
      if (int rc = K.reorth_norm(V, ldv, j, w, &bj, nullptr)) return -rc;
    } else {
      // w -= alpha_j v_j and ||w|| in one pass
      double n2;
      if (int rc = K.sub_coef(w, vj, &n2)) return -rc;
      bj = std::sqrt(n2);
    }
    const double aj = K.h_s[ALPHA_SLOT].x;
    alpha[j - 1] = aj;
    if (po) {
      if (po->update(j, alpha, beta, bj)) {
        double hm = 0;
        if (int rc = K.reorth_norm(V, ldv, j, w, &bj, nullptr, &hm)) return -rc;
        po->hmax = std::max(po->hmax, hm);
        po->reset(j);
      }
    }
    beta[j - 1] = bj;
    const double sq = 1.4901161193847656e-08;  // sqrt(eps(Float64))
    if (bj <= sq * std::max(1.0, std::fabs(aj))) {
      bool stop = true;
      if (brk >= 0) {
        // ground state with a tolerance: below sqrt(eps)|alpha| beta can be
        // cancellation in w rather than an invariant subspace, and the Ritz
        // residuals are only bounded by beta.  Orthogonalise w against the
        // whole basis and stop only when what remains is within brk.
        if (int rc = K.reorth_norm(V, ldv, j, w, &bj, nullptr)) return -rc;
        if (po) po->reset(j);
        beta[j - 1] = bj;
        stop = bj <= brk;
      }
      if (stop) {
        *m = j;
        beta[j - 1] = 0.0;
        return 1;
      }
    }
    *m = j;
    if (j + 1 <= ncols) {
      char *vn = (char *)V + (size_t)j * ldv * es;
      if (int rc = K.scale(vn, w, 1.0 / bj)) return -rc;
    }
  }
  return 0;
}

// The same steps with full reorthogonalisation and the reference's breakdown
// rule, their scalars left on the device (k_lz_step): the host queues every
// step's kernels without waiting and synchronises once at the end.  Each
// step runs the host loop's kernels on the same data in the same order, so
// alpha, beta and the basis are the host loop's bit for bit.  A step that
// wants a second Gram-Schmidt sweep is not settled here: the device steps
// from it on are discarded and the host loop redoes them (rare; a Lanczos
// vector is nearly orthogonal to the basis already).
static int lanczos_extend_dev(edk_op *op, KrylovWork &K, void *V, int64_t ldv, void *w, double *alpha,
                              double *beta, int *m, int m_target, int64_t *matvecs, int ncols) {
  const size_t es = K.esize();
  const int m0 = *m;
  hipStream_t s = K.ctx->stream;
  double2 *lz = K.d_lz;
  const int64_t n = K.n;
  const unsigned g = (unsigned)((n + 255) / 256);
  const int dtype = K.dtype;
  // fold the one-block reductions into their consumers while every block's
  // read of the partials stays small (e.g. 441 blocks x 441 partials at
  // D = 112 720)
  const bool fold = (int64_t)g * K.nblk <= (1 << 18);
  k_lz_start<<<1, 1, 0, s>>>(lz, m0, m0 >= 1 ? beta[m0 - 1] : 0.0);
  for (int j = m0 + 1; j <= m_target; j++) {
    char *vj = (char *)V + (size_t)(j - 1) * ldv * es;
    if (int rc = edk_matvec_launch(op, dtype, w, vj, 1, 0, s)) return -rc;
    // w -= beta_{j-1} v_{j-1} and alpha_j = v_j^H w, into lz[j - 1]
    if (j > 1) {
      const char *vjm1 = (const char *)V + (size_t)(j - 2) * ldv * es;
      DISPATCH(VT, k_axpy_dot_part<VT><<<K.nblk, RED_TPB, 0, s>>>((VT *)w, (const VT *)vjm1, make_double2(0, 0),
                                                                  lz + LZ_M + (j - 2), -1.0, (const VT *)vj, n,
                                                                  K.d_part));
    } else {
      DISPATCH(VT, k_dot_part<VT><<<K.nblk, RED_TPB, 0, s>>>((const VT *)vj, (const VT *)w, n, K.d_part));
    }
    if (fold) {
      DISPATCH(VT, k_axpy_re_red<VT><<<g, RED_TPB, 0, s>>>((VT *)w, (const VT *)vj, K.d_part, K.nblk, lz + (j - 1),
                                                           -1.0, n));
    } else {
      k_reduce_part<<<1, RED_TPB, 0, s>>>(K.d_part, K.nblk, lz + (j - 1));
      DISPATCH(VT, k_axpy_re_dev<VT><<<g, 256, 0, s>>>((VT *)w, (const VT *)vj, lz + (j - 1), -1.0, n));
    }
    // one sweep, then ||w||^2 and the step's scalars
    if (int rc = K.cgs_sweep(V, ldv, j, w)) return -rc;
    DISPATCH(VT, k_dot_part<VT><<<K.nblk, RED_TPB, 0, s>>>((const VT *)w, (const VT *)nullptr, n, K.d_part));
    char *vn = (char *)V + (size_t)j * ldv * es;
    const bool next = j + 1 <= ncols;
    if (fold) {
      DISPATCH(VT, k_lz_step_scale<VT><<<next ? g : 1u, RED_TPB, 0, s>>>(K.d_part, K.nblk, K.d_s, j, lz, (VT *)vn,
                                                                         (const VT *)w, n, next ? 1 : 0));
    } else {
      k_lz_step<<<1, RED_TPB, 0, s>>>(K.d_part, K.nblk, K.d_s, j, lz);
      if (next) DISPATCH(VT, k_scale_dev<VT><<<g, 256, 0, s>>>((VT *)vn, (const VT *)w, lz + 2 * LZ_M, n));
    }
    if (hipGetLastError() != hipSuccess) return -edk_fail(EDK_ERR_HIP, "Lanczos step launch");
  }
  std::vector<double2> h(2 * LZ_M + 2);
  if (hipMemcpyAsync(h.data(), lz, h.size() * sizeof(double2), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return -edk_fail(EDK_ERR_HIP, "Lanczos step scalars");
  const int stop = (int)h[2 * LZ_M + 1].x, why = (int)h[2 * LZ_M + 1].y;
  const int last = stop ? (why == 1 ? stop - 1 : stop) : m_target;
  for (int j = m0 + 1; j <= last; j++) {
    alpha[j - 1] = h[j - 1].x;
    beta[j - 1] = h[LZ_M + j - 1].x;
  }
  *matvecs += last - m0;
  *m = last;
  if (stop && why == 2) {
    beta[stop - 1] = 0.0;
    return 1;
  }
  if (stop) return lanczos_extend_host(op, K, V, ldv, w, alpha, beta, m, m_target, matvecs, ncols, nullptr, -1.0);
  return 0;
}

// EDK_KRYLOV_HOST_STEPS=1 keeps every step on the host loop (the device-driven
// steps are tested against it bit for bit)
int lanczos_extend(edk_op *op, KrylovWork &K, void *V, int64_t ldv, void *w, double *alpha, double *beta, int *m,
                   int m_target, int64_t *matvecs, int ncols, PartialOrth *po, double brk) {
  const char *env = getenv("EDK_KRYLOV_HOST_STEPS");
  const bool host_only = env && env[0] == '1';
  if (!host_only && !po && brk < 0 && m_target - *m >= 2 && K.d_lz)
    return lanczos_extend_dev(op, K, V, ldv, w, alpha, beta, m, m_target, matvecs, ncols);
  return lanczos_extend_host(op, K, V, ldv, w, alpha, beta, m, m_target, matvecs, ncols, po, brk);
}

// ---------------------------------------------------------------------------
// Lanczos ground state
// ---------------------------------------------------------------------------
int edk_lanczos_run(edk_op *op, edk_ctx *ctx, int dtype, const void *x0, int m, double tol, int refine, double *E0,
                    double *resid, int *steps, void *vec_out, int reorth, int *reorth_steps) {
  if (!op || !ctx || !x0 || !E0 || !resid || m < 1) return edk_fail(EDK_ERR_ARG, "edk_lanczos_ground: bad argument");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "bad dtype");
  if (reorth != EDK_REORTH_FULL && reorth != EDK_REORTH_PARTIAL)
    return edk_fail(EDK_ERR_ARG, "reorth must be EDK_REORTH_FULL or EDK_REORTH_PARTIAL");
  if (m > EDK_KRYLOV_MAXM) return edk_fail(EDK_ERR_UNSUPPORTED, "Lanczos dimension above 1024");
  int64_t D = op->basis->D;
  if (D < 1) return edk_fail(EDK_ERR_ARG, "empty sector");
  if (m > D) m = (int)D;
  EDK_HIP(hipSetDevice(ctx->device));
  KrylovWork K;
  if (int rc = K.init(ctx, D, dtype, m)) {
    K.release();
    return rc;
  }
  size_t es = K.esize();
  // the basis grows by doubling from 64 columns (most solves stop well
  // before m): smaller, reusable allocations instead of m + 1 columns at once
  int cap = std::min(m + 1, 64);
  void *V = nullptr, *w = nullptr;
  if (edk_dalloc(ctx, &V, (size_t)D * cap * es) != hipSuccess || edk_dalloc(ctx, &w, (size_t)D * es) != hipSuccess) {
    edk_dfree(ctx, V);
    K.release();
    return edk_fail(EDK_ERR_OOM, "Lanczos basis does not fit in device memory");
  }
  int rc = EDK_OK;
  double nrm = 0;
  std::vector<double> alpha(m, 0.0), beta(m, 0.0), lam(m), Q((size_t)m * m);
  int mm = 0, next_ql = 0;
  int64_t mv = 0;
  rc = K.norm(x0, &nrm);
  if (!rc && !(nrm > 0)) rc = edk_fail(EDK_ERR_NUMERIC, "Initial state has zero norm");
  if (!rc) rc = K.scale(V, x0, 1.0 / nrm);
  double e0 = 0, res = 0;
  // unit roundoff of the stored vectors
  PartialOrth po(dtype == EDK_C64 ? 2.220446049250313e-16 : 1.1920928955078125e-07);
  PartialOrth *pop = reorth == EDK_REORTH_PARTIAL ? &po : nullptr;
  for (int j = 1; !rc && j <= m; j++) {
    if (j + 1 > cap) {
      // step j writes column j: move columns 0..j-1 to a basis twice as wide
      const int ncap = std::min(m + 1, 2 * cap);
      void *V2 = nullptr;
      if (edk_dalloc(ctx, &V2, (size_t)D * ncap * es) != hipSuccess) {
        rc = edk_fail(EDK_ERR_OOM, "Lanczos basis does not fit in device memory");
        break;
      }
      if (hipMemcpyAsync(V2, V, (size_t)D * j * es, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) {
        edk_dfree(ctx, V2);
        rc = edk_fail(EDK_ERR_HIP, "Lanczos basis copy");
        break;
      }
      edk_dfree(ctx, V);
      V = V2;
      cap = ncap;
    }
    int br = lanczos_extend(op, K, V, D, w, alpha.data(), beta.data(), &mm, j, &mv, m + 1, pop, tol > 0 ? tol : -1.0);
    if (br < 0) {
      rc = -br;
      break;
    }
    if (br) break;
    if (tol > 0 && mm >= next_ql) {
      // residual estimate |beta_m q_m| of the lowest Ritz pair, O(m) per step
      double th, qm;
      lowest_ritz(mm, alpha.data(), beta.data(), &th, &qm);
      if (std::fabs(beta[mm - 1] * qm) <= tol) {
        // confirmed by the QL solver before stopping: when the two lowest
        // Ritz values nearly coincide, inverse iteration returns a mixture
        // of their vectors and can understate the bottom component
        const int last = mm - 1;
        if ((rc = edk_tridiag_eig(mm, alpha.data(), beta.data(), lam.data(), Q.data(), 1, &last))) break;
        if (std::fabs(beta[mm - 1] * Q[0]) <= tol) break;
        next_ql = mm + std::max(4, mm / 16);  // not confirmed: test again a few steps on
      }
    }
  }
  if (!rc && mm > 0) {
    // the reported pair: the QL solver once, at the final step
    const int last = mm - 1;
    rc = edk_tridiag_eig(mm, alpha.data(), beta.data(), lam.data(), Q.data(), 1, &last);
    if (!rc) {
      e0 = lam[0];
      res = beta[mm - 1] == 0.0 ? 0.0 : std::fabs(beta[mm - 1] * Q[0]);
    }
  }
  // Partial reorthogonalisation leaves out of T the coefficients it removes
  // (up to ~sqrt(eps) each), so the Ritz vector's true residual can sit near
  // 1e-8 whatever |beta_m q_m| says.  With a tolerance (ComplexF64), the true
  // residual is checked once the estimate converged; above tol the run
  // restarts from the Ritz vector with full reorthogonalisation for the
  // remaining steps.
  const bool check = !rc && pop && tol > 0 && dtype == EDK_C64 && mm > 0 && mm < m && res <= tol;
  double true_res = 0;
  void *x1 = nullptr;
  if (!rc && (refine || vec_out || check))
    rc = edk_tridiag_eig(mm, alpha.data(), beta.data(), lam.data(), Q.data(), mm, nullptr);
  if (!rc && (refine || vec_out || check)) {
    std::vector<double2> c(mm);
    for (int k = 0; k < mm; k++) c[k] = make_double2(Q[k], 0.0);
    // Ritz vector in FP64 storage (accumulated in FP64 from the basis)
    void *y = nullptr, *hy = nullptr;
    if (edk_dalloc(ctx, &y, (size_t)D * 16) != hipSuccess) rc = edk_fail(EDK_ERR_OOM, "Ritz vector");
    if (!rc) rc = K.combine(V, D, mm, c.data(), 1.0, y, EDK_C64);
    if (!rc && (refine || check)) {
      // FP64 Rayleigh quotient and residual with the ComplexF64 matvec
      KrylovWork K64;
      if (edk_dalloc(ctx, &hy, (size_t)D * 16) != hipSuccess) rc = edk_fail(EDK_ERR_OOM, "refinement vector");
      if (!rc) rc = K64.init(ctx, D, EDK_C64, 1, 1);
      if (!rc) rc = edk_matvec_launch(op, EDK_C64, hy, y, 1, 0, ctx->stream);
      double2 num, den;
      if (!rc) rc = K64.dot(y, hy, &num);
      if (!rc) rc = K64.dot(y, y, &den);
      if (!rc) {
        double E = num.x / den.x;
        rc = K64.axpy(hy, y, make_double2(-E, 0.0));
        double rn = 0;
        if (!rc) rc = K64.norm(hy, &rn);
        if (!rc) {
          true_res = rn / std::sqrt(den.x);
          if (refine) {
            e0 = E;
            res = true_res;
          }
        }
      }
      K64.release();
      edk_dfree(ctx, hy);
    }
    if (!rc && check && true_res > tol) {
      // the restart vector (ComplexF64, the run's dtype)
      if (edk_dalloc(ctx, &x1, (size_t)D * 16) != hipSuccess) rc = edk_fail(EDK_ERR_OOM, "restart vector");
      if (!rc && hipMemcpyAsync(x1, y, (size_t)D * 16, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
        rc = edk_fail(EDK_ERR_HIP, "restart vector copy");
    }
    if (!rc && vec_out && !x1) {
      if (dtype == EDK_C64) rc = hipMemcpyAsync(vec_out, y, (size_t)D * 16, hipMemcpyDeviceToDevice, ctx->stream) == hipSuccess ? EDK_OK : edk_fail(EDK_ERR_HIP, "copy Ritz vector");
      else {
        k_scale<float2, double2><<<(unsigned)((D + 255) / 256), 256, 0, ctx->stream>>>((float2 *)vec_out, (const double2 *)y, 1.0, D);
      }
      if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = edk_fail(EDK_ERR_HIP, "Ritz vector sync");
    }
    edk_dfree(ctx, y);
  }
  hipStreamSynchronize(ctx->stream);
  edk_dfree(ctx, V);
  edk_dfree(ctx, w);
  K.release();
  if (!rc && x1) {
    int s2 = 0, re2 = 0;
    rc = edk_lanczos_run(op, ctx, dtype, x1, m - mm, tol, refine, E0, resid, &s2, vec_out, EDK_REORTH_FULL, &re2);
    edk_dfree(ctx, x1);
    if (rc) return rc;
    if (steps) *steps = mm + s2;
    if (reorth_steps) *reorth_steps = po.reorths + re2;
    return EDK_OK;
  }
  edk_dfree(ctx, x1);
  if (rc) return rc;
  *E0 = e0;
  *resid = res;
  if (steps) *steps = mm;
  if (reorth_steps) *reorth_steps = pop ? po.reorths : mm;
  return EDK_OK;
}

extern "C" int edk_lanczos_ground(edk_op *op, int dtype, const void *x0, int m, double tol, int refine, double *E0,
                                  double *resid, int *steps, void *vec_out) {
  if (!op) return edk_fail(EDK_ERR_ARG, "edk_lanczos_ground: null operator");
  return edk_lanczos_run(op, op->ctx, dtype, x0, m, tol, refine, E0, resid, steps, vec_out);
}

extern "C" int edk_lanczos_ground_ex(edk_op *op, int dtype, const void *x0, int m, double tol, int refine, int reorth,
                                     double *E0, double *resid, int *steps, int *reorth_steps, void *vec_out) {
  if (!op) return edk_fail(EDK_ERR_ARG, "edk_lanczos_ground_ex: null operator");
  return edk_lanczos_run(op, op->ctx, dtype, x0, m, tol, refine, E0, resid, steps, vec_out, reorth, reorth_steps);
}

// ---------------------------------------------------------------------------
// KrylovEvolutionCache
// ---------------------------------------------------------------------------
struct edk_krylov {
  edk_op *H;
  edk_ctx *ctx;  // context (stream) the cache runs on
  int dtype;
  int64_t N;
  double t_anchor = 0, t_latest = 0;
  void *psi_anchor = nullptr;
  double norm_anchor = 0;
  void *V = nullptr;  // N x (m_max+1)
  void *w = nullptr;
  std::vector<double> alpha, beta, lambda, Q, Qe1, Qem;
  int m = 0;
  double tol;
  int m_init, m_max, extend_step, nsample, bisect_iters, extend_basis, normalize_output;
  edk_krylov_diag diag;
  std::vector<double> accepted;
  KrylovWork K;
  // partial reorthogonalisation (EDK_REORTH_PARTIAL): the omega estimates of
  // the current Lanczos run, restarted with the basis; null: full
  std::unique_ptr<PartialOrth> po;
  double po_eps = 0;
};

extern "C" void edk_krylov_default_params(edk_krylov_params *p) {
  p->tol = 1e-10;
  p->m_init = 30;
  p->m_max = 60;
  p->extend_step = 10;
  p->nsample = 13;
  p->bisect_iters = 12;
  p->extend_basis = 1;
  p->normalize_output = 0;
  p->reorth = EDK_REORTH_FULL;
}

// _build_reduced! (TimeEvolution.jl:311-323)
static int build_reduced(edk_krylov *c) {
  int m = c->m;
  if (m < 1) return edk_fail(EDK_ERR_NUMERIC, "Cannot build reduced propagator for empty basis");
  c->lambda.assign(m, 0);
  c->Q.assign((size_t)m * m, 0);
  c->Qe1.assign(m, 0);
  c->Qem.assign(m, 0);
  if (int rc = edk_tridiag_eig(m, c->alpha.data(), c->beta.data(), c->lambda.data(), c->Q.data(), m, nullptr))
    return rc;
  for (int k = 0; k < m; k++) {
    c->Qe1[k] = c->Q[0 + (size_t)k * m];
    c->Qem[k] = c->Q[(m - 1) + (size_t)k * m];
  }
  return EDK_OK;
}

static inline double2 cisd(double x) { return make_double2(std::cos(x), std::sin(x)); }

// _defect_monitor (TimeEvolution.jl:341-350)
static double defect(const edk_krylov *c, double tau) {
  int m = c->m;
  double bm = c->beta[m - 1];
  if (bm == 0.0) return 0.0;
  double ar = 0, ai = 0;
  for (int k = 0; k < m; k++) {
    double2 z = cisd(-tau * c->lambda[k]);
    ar += c->Qem[k] * z.x * c->Qe1[k];
    ai += c->Qem[k] * z.y * c->Qe1[k];
  }
  return std::fabs(bm) * std::hypot(ar, ai);
}

// _reconstruct_state! (TimeEvolution.jl:352-362) with _reduced_coeffs (:329-338)
static int reconstruct(edk_krylov *c, double tau, void *out) {
  int m = c->m;
  std::vector<double2> ph(m), cc(m);
  for (int k = 0; k < m; k++) {
    double2 z = cisd(-tau * c->lambda[k]);
    ph[k] = make_double2(z.x * c->Qe1[k], z.y * c->Qe1[k]);
  }
  for (int i = 0; i < m; i++) {
    double sr = 0, si = 0;
    for (int k = 0; k < m; k++) {
      sr += c->Q[i + (size_t)k * m] * ph[k].x;
      si += c->Q[i + (size_t)k * m] * ph[k].y;
    }
    cc[i] = make_double2(sr, si);
  }
  if (int rc = c->K.combine(c->V, c->N, m, cc.data(), c->norm_anchor, out, c->dtype)) return rc;
  if (c->normalize_output) {
    double nrm;
    if (int rc = c->K.norm(out, &nrm)) return rc;
    if (nrm > 0)
      if (int rc = c->K.scale(out, out, 1.0 / nrm)) return rc;
  }
  return EDK_OK;
}

// _effective_nsample (TimeEvolution.jl:383-394)
static int eff_nsample(const edk_krylov *c, double tau) {
  int base = c->nsample;
  if (c->m <= 1) return base;
  double om = c->lambda[c->m - 1] - c->lambda[0];
  if (om <= 0) return base;
  int nn = (int)std::ceil(4.0 * om * tau / M_PI) + 1;
  return std::min(std::max(nn, base), 513);
}

// _max_valid_interval (TimeEvolution.jl:396-429)
static double max_valid(const edk_krylov *c, double tau) {
  if (!(tau > 0)) return 0.0;
  if (std::fabs(c->beta[c->m - 1]) <= c->tol) return tau;
  int n = eff_nsample(c, tau);
  double prev = 0.0;
  for (int k = 1; k <= n; k++) {
    double x = tau * (double)(k - 1) / (double)(n - 1);
    double eta = defect(c, x);
    if (!std::isfinite(eta) || eta > c->tol) {
      if (k == 1) return 0.0;
      double lo = prev, hi = x;
      for (int q = 0; q < c->bisect_iters; q++) {
        double mid = 0.5 * (lo + hi);
        if (defect(c, mid) > c->tol) hi = mid;
        else lo = mid;
      }
      return lo;
    }
    prev = x;
  }
  return tau;
}

static int lanczos_build(edk_krylov *c, int m_target) {
  m_target = std::min(m_target, c->m_max);
  int r = lanczos_extend(c->H, c->K, c->V, c->N, c->w, c->alpha.data(), c->beta.data(), &c->m, m_target,
                         &c->diag.matvecs, c->m_max + 1, c->po.get());
  return r < 0 ? -r : EDK_OK;
}

// _extend_basis! (TimeEvolution.jl:434-447); *did = 1 when the basis grew
static int extend(edk_krylov *c, int *did) {
  *did = 0;
  if (!c->extend_basis || c->m >= c->m_max) return EDK_OK;
  int old = c->m;
  if (int rc = lanczos_build(c, std::min(c->m + c->extend_step, c->m_max))) return rc;
  if (c->m > old) {
    if (int rc = build_reduced(c)) return rc;
    c->diag.basis_extensions++;
    c->diag.max_dim_used = std::max<int64_t>(c->diag.max_dim_used, c->m);
    *did = 1;
  }
  return EDK_OK;
}

// _restart_from_new_anchor! (TimeEvolution.jl:449-477)
static int restart(edk_krylov *c, double tau) {
  if (int rc = reconstruct(c, tau, c->psi_anchor)) return rc;
  c->t_anchor += tau;
  double nrm;
  if (int rc = c->K.norm(c->psi_anchor, &nrm)) return rc;
  if (!(nrm > 0)) return edk_fail(EDK_ERR_NUMERIC, "Restart failed: anchor state has zero norm");
  c->norm_anchor = nrm;
  std::fill(c->alpha.begin(), c->alpha.end(), 0.0);
  std::fill(c->beta.begin(), c->beta.end(), 0.0);
  c->m = 0;
  if (c->po) c->po = std::make_unique<PartialOrth>(c->po_eps);
  if (int rc = c->K.scale(c->V, c->psi_anchor, 1.0 / nrm)) return rc;
  if (int rc = lanczos_build(c, c->m_init)) return rc;
  if (int rc = build_reduced(c)) return rc;
  c->diag.basis_builds++;
  c->diag.restarts++;
  c->diag.max_dim_used = std::max<int64_t>(c->diag.max_dim_used, c->m);
  return EDK_OK;
}

// _ensure_valid_up_to! (TimeEvolution.jl:486-520)
static int ensure_valid(edk_krylov *c, double target) {
  double rem = target;
  if (rem <= 0) return EDK_OK;
  while (true) {
    double v = max_valid(c, rem);
    if (v >= rem) return EDK_OK;
    int did = 0;
    if (c->extend_basis && c->m < c->m_max) {
      if (int rc = extend(c, &did)) return rc;
      if (did) continue;
    }
    if (v <= 0) {
      if (c->m < c->m_max) {
        if (int rc = extend(c, &did)) return rc;
        if (did) continue;
      }
      return edk_fail(EDK_ERR_NUMERIC,
                      "KrylovEvolutionCache: basis is not valid at tau=0; consider increasing m_max or loosening tol");
    }
    c->accepted.push_back(v);
    c->diag.n_accepted_intervals++;
    if (int rc = restart(c, v)) return rc;
    rem -= v;
  }
}

static void krylov_free(edk_krylov *c) {
  if (!c) return;
  edk_dfree(c->ctx, c->psi_anchor);
  edk_dfree(c->ctx, c->V);
  edk_dfree(c->ctx, c->w);
  c->K.release();
  delete c;
}

int edk_krylov_create_on(edk_op *op, edk_ctx *ctx, int dtype, const void *psi0, const edk_krylov_params *p,
                         edk_krylov **out) {
  if (!op || !ctx || !psi0 || !out) return edk_fail(EDK_ERR_ARG, "edk_krylov_create: null argument");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "bad dtype");
  edk_krylov_params prm;
  edk_krylov_default_params(&prm);
  if (p) prm = *p;
  int64_t N = op->basis->D;
  // argument checks of KrylovEvolutionCache (TimeEvolution.jl:178-185)
  if (N <= 0) return edk_fail(EDK_ERR_ARG, "Initial state must be non-empty");
  if (prm.m_init < 1) return edk_fail(EDK_ERR_ARG, "m_init must be >= 1");
  if (prm.m_max < prm.m_init) return edk_fail(EDK_ERR_ARG, "m_max must be >= m_init");
  if (prm.m_max > N) prm.m_max = (int)N;
  if (prm.m_init > prm.m_max) prm.m_init = prm.m_max;
  if (prm.nsample < 2) return edk_fail(EDK_ERR_ARG, "nsample must be >= 2");
  if (prm.bisect_iters < 0) return edk_fail(EDK_ERR_ARG, "bisect_iters must be >= 0");
  if (prm.extend_step < 1) return edk_fail(EDK_ERR_ARG, "extend_step must be >= 1");
  if (prm.m_max > EDK_KRYLOV_MAXM) return edk_fail(EDK_ERR_UNSUPPORTED, "m_max above 1024");
  if (prm.reorth != EDK_REORTH_FULL && prm.reorth != EDK_REORTH_PARTIAL)
    return edk_fail(EDK_ERR_ARG, "reorth must be EDK_REORTH_FULL or EDK_REORTH_PARTIAL");
  EDK_HIP(hipSetDevice(ctx->device));
  edk_krylov *c = new edk_krylov();
  c->H = op;
  c->ctx = ctx;
  c->dtype = dtype;
  c->N = N;
  c->tol = prm.tol;
  c->m_init = prm.m_init;
  c->m_max = prm.m_max;
  c->extend_step = prm.extend_step;
  c->nsample = prm.nsample;
  c->bisect_iters = prm.bisect_iters;
  c->extend_basis = prm.extend_basis;
  c->normalize_output = prm.normalize_output;
  c->po_eps = dtype == EDK_C64 ? 2.220446049250313e-16 : 1.1920928955078125e-07;
  if (prm.reorth == EDK_REORTH_PARTIAL) c->po = std::make_unique<PartialOrth>(c->po_eps);
  memset(&c->diag, 0, sizeof(c->diag));
  c->alpha.assign(c->m_max, 0.0);
  c->beta.assign(c->m_max, 0.0);
  int rc = c->K.init(ctx, N, dtype, c->m_max + 1);
  size_t es = c->K.esize();
  if (!rc && (edk_dalloc(ctx, &c->psi_anchor, N * es) != hipSuccess ||
              edk_dalloc(ctx, &c->V, (size_t)N * (c->m_max + 1) * es) != hipSuccess ||
              edk_dalloc(ctx, &c->w, N * es) != hipSuccess))
    rc = edk_fail(EDK_ERR_OOM, "Krylov cache does not fit in device memory");
  if (!rc && hipMemcpyAsync(c->psi_anchor, psi0, N * es, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
    rc = edk_fail(EDK_ERR_HIP, "copy psi0");
  double nrm = 0;
  if (!rc) rc = c->K.norm(c->psi_anchor, &nrm);
  if (!rc && !(nrm > 0)) rc = edk_fail(EDK_ERR_NUMERIC, "Initial state has zero norm");
  if (!rc) {
    c->norm_anchor = nrm;
    rc = c->K.scale(c->V, c->psi_anchor, 1.0 / nrm);
  }
  if (!rc) rc = lanczos_build(c, c->m_init);
  if (!rc) rc = build_reduced(c);
  if (rc) {
    krylov_free(c);
    return rc;
  }
  c->diag.basis_builds = 1;
  c->diag.max_dim_used = c->m;
  *out = c;
  return EDK_OK;
}

extern "C" int edk_krylov_create(edk_op *op, int dtype, const void *psi0, const edk_krylov_params *p,
                                 edk_krylov **out) {
  if (!op) return edk_fail(EDK_ERR_ARG, "edk_krylov_create: null argument");
  return edk_krylov_create_on(op, op->ctx, dtype, psi0, p, out);
}

extern "C" int edk_krylov_evolve(edk_krylov *c, double t, void *out) {
  if (!c || !out) return edk_fail(EDK_ERR_ARG, "edk_krylov_evolve: null argument");
  EDK_HIP(hipSetDevice(c->ctx->device));
  if (t < c->t_latest)
    return edk_fail(EDK_ERR_ARG, "timeevolve!: requested time precedes the last served time; backward evolution is not supported");
  double tau = t - c->t_anchor;
  if (tau < 0) return edk_fail(EDK_ERR_ARG, "timeevolve!: requested time precedes anchor time");
  if (int rc = ensure_valid(c, tau)) return rc;
  if (int rc = reconstruct(c, t - c->t_anchor, out)) return rc;
  c->t_latest = t;
  c->diag.total_times_served++;
  EDK_HIP(hipStreamSynchronize(c->ctx->stream));
  return EDK_OK;
}

extern "C" int edk_krylov_diagnostics(edk_krylov *c, edk_krylov_diag *d) {
  if (!c || !d) return edk_fail(EDK_ERR_ARG, "null argument");
  *d = c->diag;
  return EDK_OK;
}

extern "C" int edk_krylov_accepted_intervals(edk_krylov *c, double *out, int64_t cap, int64_t *n) {
  if (!c || !n) return edk_fail(EDK_ERR_ARG, "null argument");
  *n = (int64_t)c->accepted.size();
  for (int64_t i = 0; i < std::min<int64_t>(cap, *n); i++) out[i] = c->accepted[i];
  return EDK_OK;
}

extern "C" int edk_krylov_destroy(edk_krylov *c) {
  if (!c) return EDK_OK;
  hipSetDevice(c->ctx->device);
  hipStreamSynchronize(c->ctx->stream);
  krylov_free(c);
  return EDK_OK;
}

// ---------------------------------------------------------------------------
// device vector kernels of the C ABI
// ---------------------------------------------------------------------------
static int vec_work(edk_ctx *ctx, int dtype, int64_t n, KrylovWork &K) {
  if (!ctx || n < 0) return edk_fail(EDK_ERR_ARG, "bad context or length");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "bad dtype");
  EDK_HIP(hipSetDevice(ctx->device));
  return K.init(ctx, n, dtype, 1);
}

extern "C" int edk_vec_scale(edk_ctx *ctx, int dtype, void *y, const void *x, double s, int64_t n) {
  KrylovWork K;
  int rc = vec_work(ctx, dtype, n, K);
  if (!rc && n > 0) rc = K.scale(y, x, s);
  K.release();
  return rc;
}

extern "C" int edk_vec_axpy(edk_ctx *ctx, int dtype, void *y, const void *x, double are, double aim, int64_t n) {
  KrylovWork K;
  int rc = vec_work(ctx, dtype, n, K);
  if (!rc && n > 0) rc = K.axpy(y, x, make_double2(are, aim));
  K.release();
  return rc;
}

extern "C" int edk_vec_dot(edk_ctx *ctx, int dtype, const void *x, const void *y, int64_t n, double *re, double *im) {
  if (!re || !im) return edk_fail(EDK_ERR_ARG, "null output");
  KrylovWork K;
  int rc = vec_work(ctx, dtype, n, K);
  double2 d = make_double2(0, 0);
  if (!rc && n > 0) rc = K.dot(x, y, &d);
  K.release();
  *re = d.x;
  *im = d.y;
  return rc;
}

extern "C" int edk_vec_nrm2(edk_ctx *ctx, int dtype, const void *x, int64_t n, double *out) {
  if (!out) return edk_fail(EDK_ERR_ARG, "null output");
  KrylovWork K;
  int rc = vec_work(ctx, dtype, n, K);
  double v = 0;
  if (!rc && n > 0) rc = K.norm(x, &v);
  K.release();
  *out = v;
  return rc;
}

// ---------------------------------------------------------------------------
// Allocation-free x^H y for callers that cannot allocate device memory (no
// counterpart in the reference).  The block partials and the reduced value
// live in one file-scope device array per GPU: 1024 partials (the most
// red_blocks() ever makes) and one result, 16.4 KB held until the process
// ends.  Its device address is looked up once per device, with the device
// current, when a context is created there; kernels and copies take that
// address, never the array's host-side name.  Every call on a GPU shares the
// array, so calls that overlap on one GPU (other streams, other threads) race.
// Nothing in it needs to start at zero: the partial pass writes every slot the
// reduction reads.
// ---------------------------------------------------------------------------
#define EDK_NOALLOC_SLOTS 1025
#define EDK_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) double2 scratch[EDK_NOALLOC_SLOTS];
static std::mutex g_scratch_mu;
static double2 *g_scratch_ptr[EDK_NOALLOC_MAX_DEVICES] = {};

int edk_noalloc_setup(int device) {
  if (device < 0 || device >= EDK_NOALLOC_MAX_DEVICES) return EDK_OK;  // the entry point refuses such devices
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return EDK_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return edk_fail(EDK_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = (double2 *)p;
  return EDK_OK;
}

extern "C" int edk_vec_dot_noalloc(edk_ctx *ctx, int dtype, const void *x, const void *y, int64_t n, double *re,
                                   double *im) {
  if (!ctx || !re || !im || n < 0) return edk_fail(EDK_ERR_ARG, "edk_vec_dot_noalloc: bad argument");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "edk_vec_dot_noalloc: bad dtype");
  if (n > 0 && (!x || !y)) return edk_fail(EDK_ERR_ARG, "edk_vec_dot_noalloc: null vector");
  const int nblk = red_blocks(n);
  if (nblk + 1 > EDK_NOALLOC_SLOTS)
    return edk_fail(EDK_ERR_UNSUPPORTED, "edk_vec_dot_noalloc: more block partials than the 1025-slot scratch");
  if (ctx->device < 0 || ctx->device >= EDK_NOALLOC_MAX_DEVICES)
    return edk_fail(EDK_ERR_UNSUPPORTED, "edk_vec_dot_noalloc: device ordinal above 63");
  double2 *ptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ptr = g_scratch_ptr[ctx->device];
  }
  if (!ptr) return edk_fail(EDK_ERR_STATE, "edk_vec_dot_noalloc: scratch address not set up on this device");
  *re = 0.0;
  *im = 0.0;
  if (n == 0) return EDK_OK;
  EDK_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  DISPATCH(VT, k_dot_part<VT><<<nblk, RED_TPB, 0, s>>>((const VT *)x, (const VT *)y, n, ptr));
  k_reduce_part<<<1, RED_TPB, 0, s>>>(ptr, nblk, ptr + (EDK_NOALLOC_SLOTS - 1));
  EDK_HIP(hipGetLastError());
  double2 *h = (double2 *)ctx->pinned;
  EDK_HIP(hipMemcpyAsync(h, ptr + (EDK_NOALLOC_SLOTS - 1), sizeof(double2), hipMemcpyDeviceToHost, s));
  EDK_HIP(hipStreamSynchronize(s));
  *re = h[0].x;
  *im = h[0].y;
  return EDK_OK;
}

