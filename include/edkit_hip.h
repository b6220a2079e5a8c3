/*
 * edkit_hip.h -- C ABI of libedkit_hip.so, the MI355X (gfx950) engine behind
 * EDKit.jl's matrix-free Operator path.
 *
 * EDKit.jl has no FFI layer of its own: its boundary is Julia multiple
 * dispatch.  Each entry point below names the Julia method it stands in for
 * (file:line in the EDKit.jl v0.5.0 tree); a Julia binding calls these through
 * `ccall`, with plain pointers and sizes as declared here.
 *
 * Conventions
 *   - Every function returns an int status (EDK_OK == 0); on failure the
 *     message is in edk_last_error() (thread-local).  Nothing throws.
 *   - Handles are opaque.  The library owns every device allocation it makes;
 *     host arrays passed in are only borrowed for the duration of the call.
 *   - Representative tables cross the boundary 1-based exactly as the
 *     reference stores `b.I`; the device keeps its own packed 0-based form.
 *   - Vectors passed to the matvec and Krylov calls are DEVICE pointers
 *     (from edk_malloc or any HIP allocation on the context's device).  A
 *     block of nb vectors is row-interleaved: element (i, c) at x[i*nb + c].
 *   - Complex vectors are interleaved (re, im): EDK_C64 = 2 x double,
 *     EDK_C32 = 2 x float.
 *   - A context is not reentrant: use one context per host thread.
 */
#ifndef EDKIT_HIP_H
#define EDKIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define EDK_OK 0
#define EDK_ERR_ARG 1         /* invalid argument (reference: error()/@assert) */
#define EDK_ERR_HIP 2         /* HIP runtime error */
#define EDK_ERR_OOM 3         /* device allocation failed */
#define EDK_ERR_UNSUPPORTED 4 /* valid in the reference, not on the device */
#define EDK_ERR_STATE 5       /* call out of order / handle still in use */
#define EDK_ERR_NUMERIC 6     /* e.g. zero-norm start vector, invalid basis at tau=0 */

/* basis kinds */
#define EDK_TENSOR 0
#define EDK_PROJECTED 1
#define EDK_TRANSLATIONAL 2
#define EDK_TRANSLATION_PARITY 3
#define EDK_ABELIAN 4
#define EDK_PARITY 5
#define EDK_FLIP 6
#define EDK_PARITY_FLIP 7
#define EDK_TRANSLATION_FLIP 8

/* vector element types */
#define EDK_C64 0 /* ComplexF64 */
#define EDK_C32 1 /* ComplexF32 (matvec in FP32; Krylov scalars stay FP64) */

typedef struct edk_ctx edk_ctx;
typedef struct edk_basis edk_basis;
typedef struct edk_op edk_op;
typedef struct edk_krylov edk_krylov;

const char *edk_last_error(void);
int edk_version(void);

/* ---- context ----------------------------------------------------------- */
/* One HIP stream per context.  stream == NULL creates a private stream;
 * otherwise the caller's hipStream_t is used (e.g. torch's current stream). */
int edk_ctx_create(int device, void *stream, edk_ctx **out);
int edk_ctx_destroy(edk_ctx *ctx);
int edk_ctx_synchronize(edk_ctx *ctx);
void *edk_ctx_stream(edk_ctx *ctx);

/* device memory helpers for callers without their own allocator */
int edk_malloc(edk_ctx *ctx, size_t bytes, void **dptr);
int edk_free(edk_ctx *ctx, void *dptr);
int edk_memcpy_htod(edk_ctx *ctx, void *dst, const void *src, size_t bytes);
int edk_memcpy_dtoh(edk_ctx *ctx, void *dst, const void *src, size_t bytes);
int edk_memset(edk_ctx *ctx, void *dptr, int value, size_t bytes);

/* Diagnostics of the device block cache that serving contexts allocate from
 * (no counterpart in the reference): bytes cached for reuse, blocks owned,
 * releases of cached blocks after an allocation failure and the bytes they
 * returned, and the releases the cache refused (a block given back twice) or
 * passed through (a block it never handed out). */
typedef struct edk_cache_stats {
  int64_t cached_bytes, owned_blocks, releases, released_bytes, double_frees, foreign_frees;
} edk_cache_stats;
int edk_cache_info(int device, edk_cache_stats *out);

/* ---- bases -------------------------------------------------------------- */
/* Keyword arguments of the reference constructors.  N < 0 means `nothing`.
 *   EDK_TENSOR              TensorBasis(L, base)            AbstractBasis.jl:234-238
 *   EDK_PROJECTED           ProjectedBasis(L, N, base)      ProjectedBasis.jl:244-260
 *   EDK_TRANSLATIONAL       TranslationalBasis(L, k, N, a, base)
 *                                                           TranslationalBasis.jl:235-266
 *   EDK_TRANSLATION_PARITY  TranslationParityBasis(L, k, p, N, a, base)
 *                                                           TranslationalParityBasis.jl:141-175
 *   EDK_ABELIAN             basis(L, N, k, a, p, z, symmetries, base)
 *                                                           AbelianBasis.jl:771-847
 *   EDK_PARITY              ParityBasis(L, p, N, base)      ParityBasis.jl:76-97
 *   EDK_FLIP                FlipBasis(L, p, N, base)        FlipBasis.jl:72-95
 *   EDK_PARITY_FLIP         ParityFlipBasis(L, p, z, N, base)
 *                                                           ParityFlipBasis.jl:87-111
 *   EDK_TRANSLATION_FLIP    TranslationFlipBasis(L, k, p, N, a, base)
 *                                                           TranslationalFlipBasis.jl:122-158
 * For EDK_ABELIAN has_k/has_p/has_z tell which keywords were given; custom
 * generators come as ncustom 1-based permutations of length L (perms),
 * momenta (ks) and optional spin-inversion flags (invs, ncustom*L bytes). */
typedef struct {
  int32_t kind;
  int32_t L, base, a;
  int32_t N;
  int32_t k, p, z;
  int32_t has_k, has_p, has_z;
  int32_t ncustom;
  const int32_t *perms;
  const int32_t *ks;
  const uint8_t *invs;
} edk_basis_spec;

/* Enumerate the sector on the device (orbit-minimum test + prefix-sum
 * compaction).  Replaces the constructors listed above. */
int edk_basis_build(edk_ctx *ctx, const edk_basis_spec *spec, edk_basis **out);
/* Build from host tables the caller already has (for example a basis whose
 * predicate `f` was applied on the host): I 1-based sorted, R norms.
 * Replaces passing a prebuilt basis object into operator(...) Operator.jl:56. */
int edk_basis_upload(edk_ctx *ctx, const edk_basis_spec *spec,
                     const int64_t *I_1based, const double *R, int64_t D,
                     edk_basis **out);
int64_t edk_basis_dim(const edk_basis *b);          /* size(b, 1) AbstractBasis.jl:142 */
int edk_basis_order(const edk_basis *b);            /* order(b) */
/* Copy the tables back: I 1-based (b.I), R (b.R).  Either may be NULL. */
int edk_basis_download(edk_basis *b, int64_t *I_1based, double *R);
/* index(b, dgt) for n states given as 0-based contents (the reference's
 * index(dgt) - 1): 1-based pos and coefficient; absent gives (0, pos 1).
 * TranslationalBasis.jl:310-341, TranslationalParityBasis.jl:192-234,
 * AbelianBasis.jl:669-693, ProjectedBasis.jl:52-57, AbstractBasis.jl:246. */
int edk_basis_index(edk_basis *b, const int64_t *contents, int64_t n,
                    int64_t *pos_1based, double *coef_re, double *coef_im);
int edk_basis_destroy(edk_basis *b);

/* ---- operators ------------------------------------------------------------ */
/* Operator{Tv,Tb}(M::Vector{SparseMatrixCSC}, I::Vector{Vector{Int}}, B)
 * (Operator.jl:24-28, built by operator(...) :56-78 / trans_inv_operator
 * :102-108), flattened:
 *   site_off[nterms+1], sites[]        1-based site lists, first site = MSB
 *   colptr_off[nterms]                 start of term t's colptr in colptr[]
 *   colptr[]                           SparseMatrixCSC.colptr, 1-based
 *   nz_off[nterms+1], rowval[]         SparseMatrixCSC.rowval, 1-based
 *   nzval[]                            complex interleaved (re, im)
 * Precondition of the gather kernels: the operator commutes with the basis'
 * symmetry group (as every reference sector test assumes). */
int edk_operator_create(edk_ctx *ctx, edk_basis *b, int64_t nterms,
                        const int64_t *site_off, const int64_t *sites,
                        const int64_t *colptr_off, const int64_t *colptr,
                        const int64_t *nz_off, const int64_t *rowval,
                        const double *nzval, edk_op **out);
int edk_operator_destroy(edk_op *op);
/* Exact number of off-diagonal local nonzeros whose target lies in the
 * sector, summed over rows (device counting pass). */
int edk_operator_count_offdiag(edk_op *op, int64_t *nnz_off);

/* y = (accumulate ? y : 0) + H x, on device pointers, nb row-interleaved
 * columns.  accumulate=1 is mul!(y, H, x) (Operator.jl:309-328); accumulate=0
 * is the zero-filled target of H*x / _apply! (:372-394, TimeEvolution.jl:237-241). */
int edk_matvec(edk_op *op, int dtype, void *y, const void *x, int64_t nb,
               int accumulate);
/* The literal host-array mul!(y, H, x): copies in, applies, copies out. */
int edk_matvec_host(edk_op *op, int dtype, void *y_host, const void *x_host,
                    int64_t nb, int accumulate);
/* ys[c] = H_c xs[c] (zero-filled targets, as H*x) for n <= 8 operators on one
 * basis that share a two-site family structure and differ in their values
 * (e.g. XXZ chains of different Delta, J1-J2 rings of different J2): one
 * launch on ops[0]'s context stream, every target canonicalised and looked up
 * once for all n columns; the same bits per column as a launch with that
 * column alone.  The loop `for c: mul!(ys[c], H_c, xs[c])` over Operator.jl:
 * 309-315 with one traversal.  EDK_ERR_UNSUPPORTED when the operators have no
 * shared specialised kernel (then call edk_matvec per operator). */
int edk_matvec_multi(edk_op *const *ops, int n, int dtype, void *const *ys, const void *const *xs);

/* ---- explicit sector matrix --------------------------------------------------- */
/* sparse(opt) / Array(opt) (Operator.jl:171-217) and the sparse!(opt) cache
 * (:221-285): assemble the sector matrix on the device, row by row, with the
 * gather kernel's matrix elements; each row sorted by column, duplicates
 * summed, exact zeros dropped.  Once assembled (and until released, the
 * reference's clear_sparse_cache!), edk_matvec with nb > 1 uses it (CSR SpMM);
 * a single vector keeps the matrix-free kernels, as in the reference.
 * edk_operator_assemble returns the number of stored entries. */
int edk_operator_assemble(edk_op *op, int64_t *nnz);
/* CSR of the sector matrix, 1-based: rowptr[D+1], col[nnz], val[2*nnz]
 * (complex interleaved).  As SparseMatrixCSC(D, D, rowptr, col, val) this is
 * the transpose; the binding takes copy(transpose(...)). */
int edk_operator_download_csr(edk_op *op, int64_t *rowptr_1based, int64_t *col_1based, double *val);
int edk_operator_release_csr(edk_op *op);

/* ---- schedules and batching ------------------------------------------------ */
/* No counterpart in the reference (its mul! has one CPU loop).  Schedules of
 * the specialised gather kernel, chosen per operator, dtype and block width:
 *   EDK_VARIANT_AUTO     library default (tile where it applies)
 *   EDK_VARIANT_GENERIC  the generic kernel (any basis and operator)
 *   EDK_VARIANT_TILE     one workgroup per chunk of whole rows (<= 1024 targets),
 *                        targets spread over all lanes; blocks of 2/4/8 vectors
 *   EDK_VARIANT_LANE2/4/8  one lane per row, 2/4/8 targets in flight (nb = 1) */
#define EDK_VARIANT_AUTO 0
#define EDK_VARIANT_GENERIC 1
#define EDK_VARIANT_TILE 2
#define EDK_VARIANT_LANE2 3
#define EDK_VARIANT_LANE4 4
#define EDK_VARIANT_LANE8 5
int edk_operator_set_variant(edk_op *op, int dtype, int64_t nb, int variant);
/* Time each applicable schedule (reps launches on scratch vectors) and keep
 * the fastest for later edk_matvec calls with this dtype and nb. */
int edk_operator_autotune(edk_op *op, int dtype, int64_t nb, int reps, int *variant, double *ms);

/* Many sectors and block widths in one launch: the matvecs of n items are
 * recorded once as a HIP graph (y = (accumulate ? y : 0) + H x per item,
 * replacing a loop of mul!(Y_s, H_s, X_s) over sectors, Operator.jl:322-328).
 * Items may use different operators/bases of one context; with streams > 1
 * they are spread over that many graph branches, balanced by D * nb, so
 * independent sectors run concurrently (streams <= 0: one branch when every
 * item has D * nb >= 2^20, else 8).  The graph keeps the pointers: the
 * vectors must stay allocated while the batch is used. */
typedef struct {
  edk_op *op;
  const void *x; /* device, D * nb row-interleaved */
  void *y;       /* device, D * nb row-interleaved, must not alias x */
  int64_t nb;
  int32_t accumulate;
} edk_batch_item;
typedef struct edk_batch edk_batch;
int edk_batch_create(edk_ctx *ctx, int dtype, const edk_batch_item *items, int64_t n, int streams,
                     edk_batch **out);
/* The same with flags: EDK_BATCH_FUSE (edk_batch_create's default) runs the
 * block products of sectors that take the same specialised block kernel as
 * one launch over all their chunks (up to 12 sectors a launch); without it
 * every item keeps its own launch.  No counterpart in the reference (mul(H, M) per sector,
 * Operator.jl:354-370, is the unit a batch is made of). */
#define EDK_BATCH_FUSE 1
int edk_batch_create_ex(edk_ctx *ctx, int dtype, const edk_batch_item *items, int64_t n, int streams, int flags,
                        edk_batch **out);
int64_t edk_batch_launches(const edk_batch *b); /* kernel launches the batch's graph holds */
int edk_batch_launch(edk_batch *b); /* one graph launch on the context stream, asynchronous */
int edk_batch_destroy(edk_batch *b);

/* ---- device vector kernels ------------------------------------------------- */
/* The BLAS-1 steps the Krylov loops run inline in the reference
 * (TimeEvolution.jl:261-303: w -= beta v, dot, w -= alpha v, norm, w / beta),
 * on device vectors of length n (dtype EDK_C64 or EDK_C32).  Reductions are
 * FP64 with a fixed two-stage order: results are bitwise reproducible. */
int edk_vec_scale(edk_ctx *ctx, int dtype, void *y, const void *x, double s, int64_t n);  /* y = s x */
int edk_vec_axpy(edk_ctx *ctx, int dtype, void *y, const void *x, double a_re, double a_im,
                 int64_t n);                                                              /* y += a x */
int edk_vec_dot(edk_ctx *ctx, int dtype, const void *x, const void *y, int64_t n, double *re,
                double *im);                                                              /* x^H y */
int edk_vec_nrm2(edk_ctx *ctx, int dtype, const void *x, int64_t n, double *out);         /* ||x|| */
/* x^H y like edk_vec_dot, allocating nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc).  The block partials live in a library-
 * owned 1025-slot (16.4 KB) device array per GPU, so any n >= 0 fits; calls
 * that overlap on one GPU (several streams or threads) race on it.  Same
 * kernels and reduction order as edk_vec_dot: bitwise the same result. */
int edk_vec_dot_noalloc(edk_ctx *ctx, int dtype, const void *x, const void *y, int64_t n, double *re,
                        double *im);

/* ---- Schmidt decomposition of a sector state ------------------------------- */
/* schmidt(v, Ainds, b) (Schmidt.jl:186-330, AbelianBasis.jl:704-719): the
 * dense Schmidt matrix of the device vector v (D elements of dtype) for the
 * subsystem of the nA 1-based sites Ainds, written to the device buffer M of
 * base^nA x base^(L-nA) ComplexF64 entries, column-major (M[ia + dA ib], ia and
 * ib the 0-based indices of the A and B digits; the B sites in increasing
 * order).  Needs a device-built basis (or a tensor/projected one). */
int edk_schmidt(edk_basis *b, int dtype, const void *v, int nA, const int32_t *Ainds, void *M);
/* The squared Schmidt values of the same matrix (svdvals(schmidt(...)) .^ 2,
 * the argument of entropy in ent_S, Schmidt.jl:150-154) in descending order:
 * s2 receives min(dA, dB) values (host array), *n their number. */
int edk_ent_spec(edk_basis *b, int dtype, const void *v, int nA, const int32_t *Ainds, double *s2,
                 int64_t *n);

/* ---- Krylov --------------------------------------------------------------- */
/* Lanczos ground state (no counterpart in the reference; built from the
 * _lanczos_build! recurrence TimeEvolution.jl:250-306 with full
 * reorthogonalisation).  x0: device vector (dtype); m: max steps; tol: stop
 * when |beta_j Q[j,1]| <= tol, confirmed by the QL solver (tol <= 0 runs m
 * steps; with tol > 0 a beta below sqrt(eps)|alpha| ends the run only when,
 * after full orthogonalisation, it is <= tol); refine: 1 adds a final FP64
 * Rayleigh quotient/residual with the ComplexF64 matvec on the Ritz vector.
 * Outputs: E0, residual, steps (all restarts included); vec_out (device,
 * dtype) optional. */
int edk_lanczos_ground(edk_op *op, int dtype, const void *x0, int m,
                       double tol, int refine, double *E0, double *resid,
                       int *steps, void *vec_out);
#define EDK_REORTH_FULL 0    /* every step: against all previous vectors (DGKS) */
#define EDK_REORTH_PARTIAL 1 /* Simon's partial reorthogonalisation: only when the
                                omega-recurrence estimate of the loss of
                                orthogonality passes sqrt(eps) (then at that step
                                and the next); the basis stays semi-orthogonal */
/* edk_lanczos_ground with a choice of reorthogonalisation (no counterpart in
 * the reference).  reorth_steps (optional): steps that reorthogonalised.
 * EDK_REORTH_PARTIAL with tol > 0 (ComplexF64) checks the true residual of
 * the Ritz vector when the estimate converged and, above tol, restarts from
 * it with full reorthogonalisation for the remaining steps. */
int edk_lanczos_ground_ex(edk_op *op, int dtype, const void *x0, int m,
                          double tol, int refine, int reorth, double *E0,
                          double *resid, int *steps, int *reorth_steps,
                          void *vec_out);

/* KrylovEvolutionCache (TimeEvolution.jl:84-218), device-resident. */
typedef struct {
  double tol;          /* 1e-10 */
  int32_t m_init;      /* 30 */
  int32_t m_max;       /* 60 */
  int32_t extend_step; /* 10 */
  int32_t nsample;     /* 13 */
  int32_t bisect_iters;/* 12 */
  int32_t extend_basis;     /* 1 */
  int32_t normalize_output; /* 0 */
  int32_t reorth;           /* EDK_REORTH_FULL (the reference's _lanczos_build!) or EDK_REORTH_PARTIAL */
} edk_krylov_params;

typedef struct {
  int64_t basis_builds, basis_extensions, restarts, matvecs,
      total_times_served, max_dim_used, n_accepted_intervals;
} edk_krylov_diag;

void edk_krylov_default_params(edk_krylov_params *p);
/* KrylovEvolutionCache(H, psi0; kwargs...) TimeEvolution.jl:166-218 */
int edk_krylov_create(edk_op *op, int dtype, const void *psi0,
                      const edk_krylov_params *params, edk_krylov **out);
/* timeevolve!(cache, t) TimeEvolution.jl:586-597; out: device vector */
int edk_krylov_evolve(edk_krylov *c, double t, void *out);
/* KrylovEvolutionDiagnostics TimeEvolution.jl:37-59 */
int edk_krylov_diagnostics(edk_krylov *c, edk_krylov_diag *d);
/* accepted_intervals (up to cap entries); returns count in *n */
int edk_krylov_accepted_intervals(edk_krylov *c, double *out, int64_t cap,
                                  int64_t *n);
int edk_krylov_destroy(edk_krylov *c);

/* ---- DoubleBasis ---------------------------------------------------------- */
/* Operator(M, I, DoubleBasis(B1, B2)) (LinearMap.jl:88-126 with
 * Operator.jl:24-78): the rectangular map from the source sector B2 to the
 * target sector B1 (size(B1) x size(B2)).  Same term arrays as
 * edk_operator_create; both bases must have the same L and base
 * (check_doublebasis_compatible, LinearMap.jl:39-43) and device-built tables. */
typedef struct edk_dop edk_dop;
int edk_double_operator_create(edk_ctx *ctx, edk_basis *b1, edk_basis *b2, int64_t nterms,
                               const int64_t *site_off, const int64_t *sites,
                               const int64_t *colptr_off, const int64_t *colptr,
                               const int64_t *nz_off, const int64_t *rowval,
                               const double *nzval, edk_dop **out);
/* mul!(y, H, x) / H * x on a DoubleBasis (Operator.jl:309-394): device
 * vectors, y size(B1) x nb, x size(B2) x nb, columns row-interleaved */
int edk_double_matvec(edk_dop *op, int dtype, void *y, const void *x, int64_t nb, int accumulate);
int edk_double_operator_destroy(edk_dop *op);
/* (B::DoubleBasis)(v) (LinearMap.jl:144-172): the basis-overlap map of
 * device vectors v (size(B2) x nb) into out (size(B1) x nb) */
int edk_double_map(edk_basis *b1, edk_basis *b2, int dtype, const void *v, void *out, int64_t nb);

/* ---- Lindblad ------------------------------------------------------------- */
/* LiouvillianMap(H, jumps) (LindbladEvolution.jl:62-92): the many-body
 * Lindbladian on vec(rho), rho a dense d x d density matrix.  H and the nj
 * jump operators: device complex128 column-major d x d matrices, the jumps
 * stored one after another (nj d^2 elements); both are copied. */
typedef struct edk_liouvillian edk_liouvillian;
int edk_liouvillian_create(edk_ctx *ctx, int64_t d, const void *H, int nj, const void *jumps,
                           edk_liouvillian **out);
/* LiouvillianMap(sparse(H), sparse.(jumps)): host CSR arrays, 0-based, complex
 * values interleaved (re, im).  H: rows h_rp[0..d] into h_col/h_val; jump q:
 * rows j_rp[q(d+1) .. q(d+1)+d] into j_col/j_val.  The products run as sparse
 * row/column passes over the dense density matrix (O(nnz d) per apply). */
int edk_liouvillian_create_csr(edk_ctx *ctx, int64_t d, const int64_t *h_rp, const int32_t *h_col,
                               const double *h_val, int nj, const int64_t *j_rp, const int32_t *j_col,
                               const double *j_val, edk_liouvillian **out);
/* A * rho (_liouvillian_apply!, LindbladEvolution.jl:114-151): y = L(x),
 * x and y distinct device d x d buffers */
int edk_liouvillian_apply(edk_liouvillian *A, const void *x, void *y);
int edk_liouvillian_destroy(edk_liouvillian *A);

/* LindbladArnoldiCache keywords (LindbladEvolution.jl:154-174, 216-228) */
typedef struct {
  double tol;               /* 1e-10 */
  int32_t m_init;           /* 20 */
  int32_t m_max;            /* 50 */
  int32_t extend_step;      /* 10 */
  int32_t nsample;          /* 17 */
  int32_t bisect_iters;     /* 12 */
  int32_t extend_basis;     /* 1 */
  int32_t normalize_trace;  /* 0 */
  int32_t record_min_eig;   /* 0 */
} edk_lindblad_params;
/* LindbladArnoldiDiagnostics counters (LindbladEvolution.jl:17-37); the
 * per-output vectors come from edk_lindblad_records */
typedef struct {
  int64_t basis_builds, basis_extensions, restarts, liouvillian_applies,
      total_times_served, max_dim_used, n_accepted_intervals;
  double max_trace_drift, max_hermiticity_drift, min_min_hermitian_eig;
} edk_lindblad_diag;
typedef struct edk_lindblad edk_lindblad;
void edk_lindblad_default_params(edk_lindblad_params *p);
/* LindbladArnoldiCache(A, dm; kwargs...) (LindbladEvolution.jl:216-273);
 * rho0: device d x d complex128, column-major.  The cache keeps A. */
int edk_lindblad_create(edk_liouvillian *A, const void *rho0, const edk_lindblad_params *params,
                        edk_lindblad **out);
/* lindblad_timeevolve!(cache, t) (LindbladEvolution.jl:598-611); out: device d x d */
int edk_lindblad_evolve(edk_lindblad *c, double t, void *out);
int edk_lindblad_diagnostics(edk_lindblad *c, edk_lindblad_diag *d);
/* accepted_intervals, trace_drifts, hermiticity_drifts and min_hermitian_eigs
 * (up to cap entries each; any pointer may be null); their lengths in *n_* */
int edk_lindblad_records(edk_lindblad *c, double *accepted, double *trace_drifts, double *herm_drifts,
                         double *min_eigs, int64_t cap, int64_t *n_accepted, int64_t *n_outputs,
                         int64_t *n_eigs);
/* current Arnoldi dimension (cache.m) */
int edk_lindblad_dim(edk_lindblad *c, int *m);
int edk_lindblad_destroy(edk_lindblad *c);

/* ---- serving --------------------------------------------------------------- */
/* No counterpart in the reference.  A server runs Lanczos ground-state and
 * Krylov time-evolution jobs (any sector and operator on its device) on
 * `workers` threads, each with its own stream, so jobs overlap on the GPU.
 * A job is admitted as a worker frees up and its device footprint fits beside
 * the running ones within hbm_budget bytes (0: 80 % of the free memory at
 * creation).  Jobs whose operators share a basis and a two-site family
 * structure (one sector, different coefficients) form a group; with batching
 * on, their matvecs are launched together step by step, one block launch for
 * up to 8 members, and admission prefers queued jobs that join a running
 * group (continuous batching).  The start vector is copied at submit; the
 * job's operator and basis must stay alive until it finishes.  A job's result
 * is independent of what else runs: the same bits alone or in any mix. */
typedef struct {
  int32_t batch;        /* 1: batch the matvecs of group members */
  double batch_wait_us; /* longest wait for a group's other members before a partial block launches */
  int32_t affinity;     /* 1: admit queued jobs joining a running group before the queue head */
} edk_server_params;
void edk_server_default_params(edk_server_params *p); /* batch 1, 250 us, affinity 1 */
#define EDK_JOB_GROUND 0 /* edk_lanczos_ground on the start vector */
#define EDK_JOB_EVOLVE 1 /* KrylovEvolutionCache(H, psi0) then timeevolve!(cache, t) */
typedef struct {
  int32_t kind;        /* EDK_JOB_GROUND or EDK_JOB_EVOLVE */
  int32_t dtype;       /* EDK_C64 or EDK_C32 */
  edk_op *op;
  const void *psi0;    /* HOST start vector (D elements of dtype); NULL: the unit
                          vector at psi0_index, i.e. productstate (ToolKit.jl:103-109) */
  int64_t psi0_index;  /* 0-based position for the product state */
  int32_t m;           /* ground: Lanczos steps at most */
  int32_t refine;      /* ground: FP64 refinement of the Ritz pair */
  double tol;          /* ground: stop at |beta_m Q[m,1]| <= tol; evolve: Krylov tol (<= 0: 1e-10) */
  double t;            /* evolve: final time */
  int32_t m_init;      /* evolve: Krylov dimensions (<= 0: 30 and 60) */
  int32_t m_max;
  int32_t want_vector; /* 1: keep the result vector (ground state or psi(t)) */
  int32_t reorth;      /* EDK_REORTH_FULL or EDK_REORTH_PARTIAL (ground state and evolution) */
} edk_job;
typedef struct {
  int32_t status;      /* EDK_OK or the error code of the job */
  int32_t steps;       /* ground: Lanczos steps; evolve: restarts */
  double E0, resid;    /* ground */
  int64_t matvecs;
  double overlap_re, overlap_im, norm; /* evolve: <psi0|psi(t)>, ||psi(t)|| */
  double t_submit, t_start, t_end;     /* seconds on the server's monotonic clock */
  char message[160];   /* error text when status != EDK_OK */
} edk_job_result;
typedef struct {
  int64_t completed, queued, running, peak_running;
  int64_t inflight_bytes, peak_inflight_bytes, budget_bytes;
} edk_server_info;
typedef struct edk_server edk_server;
int edk_server_create(int device, int workers, size_t hbm_budget, edk_server **out); /* default params */
int edk_server_create_ex(int device, int workers, size_t hbm_budget, const edk_server_params *p,
                         edk_server **out);
int edk_server_submit(edk_server *s, const edk_job *job, int64_t *ticket);
/* Wait for a job (timeout_s < 0: no limit; EDK_ERR_STATE on timeout) and
 * release its ticket; vec_host (D elements) receives the kept vector. */
int edk_server_wait(edk_server *s, int64_t ticket, double timeout_s, edk_job_result *res,
                    void *vec_host);
double edk_server_clock(edk_server *s); /* the server's clock, seconds */
int edk_server_stats(edk_server *s, edk_server_info *st);
/* block launches of batched matvecs so far, the columns they carried, and
 * the columns per launch weighted by the sector dimension */
int edk_server_batch_stats(edk_server *s, int64_t *launches, int64_t *columns, double *dim_weighted);
int edk_server_destroy(edk_server *s); /* queued jobs are dropped, running ones finish */

#ifdef __cplusplus
}
#endif
#endif
