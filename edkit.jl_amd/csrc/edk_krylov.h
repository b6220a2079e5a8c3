// edk_krylov.h -- device BLAS-1 / reorthogonalisation helpers shared by the
// Lanczos and Krylov-evolution drivers.
#pragma once
#include "edk_matvec.h"

#define EDK_KRYLOV_MAXM 1024
// slot of d_s / h_s that holds a norm while d_s[0..j) holds coefficients
#define NORM_SLOT (EDK_KRYLOV_MAXM + 1)
// host slot (h_s only) receiving a Lanczos coefficient left on the device
#define ALPHA_SLOT (EDK_KRYLOV_MAXM + 2)

// Deterministic FP64 vector kernels over length-n vectors of one dtype.
struct KrylovWork {
  edk_ctx *ctx = nullptr;
  int64_t n = 0;
  int dtype = EDK_C64;
  int nblk = 1;
  double2 *d_part = nullptr;  // block partials
  double2 *d_s = nullptr;     // reduced coefficients
  // device-driven Lanczos steps: alpha_j in [0, M), (beta_j, second sweep
  // wanted) in [M, 2M), 1/beta of the last step in [2M], (first step the
  // device loop could not settle, reason) in [2M + 1]; M = EDK_KRYLOV_MAXM
  double2 *d_lz = nullptr;
  double2 *h_s = nullptr;     // pinned host mirror
  // slot: which of the context's two pinned scalar buffers to use
  int init(edk_ctx *c, int64_t n, int dtype, int maxm, int slot = 0);
  void release();
  size_t esize() const;
  int dot(const void *a, const void *b, double2 *out);  // sum conj(a) b (b == null: |a|^2)
  int norm(const void *a, double *out);
  // w += alpha x, then out = y^H w over the updated w (y == null: out.x = |w|^2), one pass
  int axpy_dot(void *w, const void *x, double2 alpha, const void *y, double2 *out);
  // conj(y) . (w += alpha x) (x == null: conj(y) . w) left in d_s[0], copied to h_s[ALPHA_SLOT]
  int coef_async(void *w, const void *x, double2 alpha, const void *y);
  // w -= Re(d_s[0]) x; with norm2_out, |w|^2 in the same pass (synchronised)
  int sub_coef(void *w, const void *x, double *norm2_out);
  int axpy(void *y, const void *x, double2 alpha);
  int scale(void *y, const void *x, double s);
  int cgs_sweep(const void *V, int64_t ldv, int j, void *w);
  // reorthogonalise w against V[:, :j] (one or two sweeps) and return ||w||
  int reorth_norm(const void *V, int64_t ldv, int j, void *w, double *wnorm, int *sweeps, double *hmax = nullptr);
  int combine(const void *V, int64_t ldv, int j, const double2 *c, double scale, void *out, int out_dtype);
};

int edk_tridiag_eig(int n, const double *alpha, const double *beta, double *d, double *Z, int nr,
                    const int *rows);
// Partial reorthogonalisation (H. D. Simon, "The Lanczos algorithm with
// partial reorthogonalization", Math. Comp. 42 (1984) 115-142): the estimates
// omega[j][k] ~ v_j^H v_k follow the Lanczos recurrence itself,
//   beta_j w_{j+1,k} = beta_k w_{j,k+1} + (alpha_k - alpha_j) w_{j,k}
//                      + beta_{k-1} w_{j,k-1} - beta_{j-1} w_{j-1,k} +- theta,
// with theta = 10 eps ||T_j|| the rounding term (||T_j|| by Gershgorin).  When
// an estimate passes sqrt(eps), the new vector and the next one are
// reorthogonalised against the whole basis and their estimates reset to eps;
// the basis stays semi-orthogonal, which keeps Ritz values and the residual
// estimate accurate to working precision (Simon, section 5).
struct PartialOrth {
  double eps = 0, anorm = 0;
  std::vector<double> wold, wcur;  // rows j-1 and j of omega (k = 1..)
  bool force = false, forced_now = false;
  int reorths = 0;
  double hmax = 0;                 // largest |v_k^H w| / ||w|| met at a reorthogonalisation
  explicit PartialOrth(double e) : eps(e), wcur(1, 1.0) {}
  // omega row j+1 from the provisional beta_j; true: reorthogonalise v_{j+1}
  bool update(int j, const double *alpha, const double *beta, double bj);
  // after v_{j+1} has been reorthogonalised against v_1..v_j
  void reset(int j);
};

// brk < 0: the reference's breakdown rule beta <= sqrt(eps) max(1, |alpha|)
// (TimeEvolution.jl:284-292).  brk >= 0 (ground state with a tolerance): a
// beta below that bound is first orthogonalised against the whole basis and
// counts as a breakdown only when the remainder is <= brk.
int lanczos_extend(edk_op *op, KrylovWork &K, void *V, int64_t ldv, void *w, double *alpha, double *beta, int *m,
                   int m_target, int64_t *matvecs, int ncols, PartialOrth *po = nullptr, double brk = -1.0);

// the Lanczos ground state and the Krylov cache on a given context (its stream
// and allocation mode), for callers that run several at once
int edk_lanczos_run(edk_op *op, edk_ctx *ctx, int dtype, const void *x0, int m, double tol, int refine, double *E0,
                    double *resid, int *steps, void *vec_out, int reorth = EDK_REORTH_FULL,
                    int *reorth_steps = nullptr);
int edk_krylov_create_on(edk_op *op, edk_ctx *ctx, int dtype, const void *psi0, const edk_krylov_params *p,
                         edk_krylov **out);
