// edk_matvec.h -- operator tables and matvec kernel arguments.
#pragma once
#include <atomic>

#include "edk_internal.h"

#define EDK_TDESC (2 + EDK_MAXSITES)  // per term: nsites, matrix id, field shifts

struct TermsDev {
  int nterms;
  const int *tdesc;        // EDK_TDESC ints per term
  const uint64_t *tpw;     // base^(L-site) per term site (content weights)
  const int *mrow_off;     // per distinct matrix: offset of its row pointers
  const int *rowptr;       // CSR row pointers into col/val
  const int *col;          // local column c of M[r, c]
  const double2 *val;      // M[r, c]
  int nmo, nrp, nnz;       // lengths of mrow_off, rowptr, col / val
};

struct MatvecArgs {
  CanonDev C;
  SearchDev S;
  TermsDev T;
  const double *R;   // basis norms b.R
  int kind;
  int64_t D;
  int L, base, w;
  int orbit_norms;   // 1: norms follow from the orbit (library-built table)
};

struct edk_op {
  edk_ctx *ctx;
  edk_basis *basis;
  bool all_real;
  int nterms;
  TermsDev terms;
  int *d_tdesc = nullptr;
  uint64_t *d_tpw = nullptr;
  int *d_mrow_off = nullptr, *d_rowptr = nullptr, *d_col = nullptr;
  double2 *d_val = nullptr;
  void *d_scratch = nullptr;
  void *fast = nullptr;       // state of the specialised kernel (edk_fast.hip)
  void *csr = nullptr;        // assembled sector matrix (edk_assembly.hip), or null
  bool disable_fast = false;  // force the generic kernel (tests compare both)
  uint32_t *d_gchunk = nullptr;  // generic tile kernel: first row of each chunk
  uint32_t ngchunk = 0;
  std::atomic<int> gstate{0};     // generic tile setup: 0 not yet, 1 built, -1 not applicable
  std::atomic<int> fstate{0};     // specialised kernel setup: 0 not yet, 1 built, -1 not applicable
  void *perm = nullptr;           // state of the Abelian two-site kernel (edk_perm.hip)
  std::atomic<int> pstate{0};     // its setup: 0 not yet, 1 built, -1 not applicable
  size_t scratch_bytes = 0;
  std::vector<int> h_tdesc, h_mrow_off, h_rowptr, h_col;
  std::vector<double2> h_val;
  MatvecArgs args() const;
};

__device__ __forceinline__ uint64_t edk_content_to_packed_d(uint64_t c, int L, int base, int w) {
  if (base == 2) return c;
  uint64_t s = 0;
  for (int q = 0; q < L; q++) {
    s |= (c % (uint64_t)base) << (q * w);
    c /= (uint64_t)base;
  }
  return s;
}

// Specialised kernels; returns true when it handled the call.
void edk_fast_release(edk_op *op);
// Abelian bases with two-site families (edk_perm.hip): one vector; false
// when it does not apply or a schedule other than the tile form is chosen
bool edk_matvec_perm(edk_op *op, int dtype, void *y, const void *x, int accumulate, hipStream_t st);
bool edk_perm_prepare(edk_op *op);  // one-time setup (allocates; run before capture)
void edk_perm_release(edk_op *op);
// assembled sector matrix: SpMM for nb > 1 (false: not assembled / not applicable)
void edk_csr_release(edk_op *op);
bool edk_matvec_csr(edk_op *op, int dtype, void *y, const void *x, int64_t nb, int accumulate, hipStream_t st);
bool edk_matvec_fast(edk_op *op, int dtype, void *y, const void *x, int64_t nb, int accumulate, hipStream_t st);
// schedule choice per (dtype, block width): forced, or timed by the autotuner
int edk_fast_set_variant(edk_op *op, int dtype, int64_t nb, int variant);
int edk_fast_autotune(edk_op *op, int dtype, int64_t nb, int reps, int *variant, double *ms);
// one-time setup of the specialised kernels (allocates; run before capture)
bool edk_matvec_prepare(edk_op *op);
// one-time setup of the generic tile kernel (allocates; run before capture)
bool edk_generic_setup(edk_op *op);
// Batched products of operators sharing a basis and a two-site family
// structure (different values): edk_fast_signature gives the structure's hash
// (false: no specialised kernel for op); edk_fast_matvec_mc sets
// ys[c] = H_c xs[c] for n <= 8 operators through one launch, with the
// row-interleaved scratch block Xblk of D x 8 elements.
bool edk_fast_signature(edk_op *op, uint64_t *sig);
int edk_fast_matvec_mc(edk_op *const *ops, int n, int dtype, const void *const *xs, void *const *ys, void *Xblk,
                       hipStream_t st);
// Matvec interception for the calling thread (the server's batcher): when
// set, edk_matvec_launch hands single-vector, non-accumulating products of
// ComplexF64/ComplexF32 vectors to it; it returns EDK_OK when it took the
// product, or a negative value to run the product directly.
struct edk_mv_hook {
  virtual int launch(edk_op *op, int dtype, void *y, const void *x, hipStream_t st) = 0;
  virtual ~edk_mv_hook() = default;
};
void edk_set_mv_hook(edk_mv_hook *h);
// y = (accumulate ? y : 0) + H x on stream st (no allocation, no synchronisation)
int edk_matvec_launch(edk_op *op, int dtype, void *y, const void *x, int64_t nb, int accumulate, hipStream_t st);
inline int edk_matvec_launch(edk_op *op, int dtype, void *y, const void *x, int64_t nb, int accumulate) {
  return edk_matvec_launch(op, dtype, y, x, nb, accumulate, op->ctx->stream);
}
