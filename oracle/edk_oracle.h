/*
 * edk_oracle.h -- CPU restatement of EDKit.jl's matrix-free Operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against (and the timed CPU baseline in bench.py).  Nothing
 * in the product library links, loads or calls it.
 *
 * Every function follows a reference function, cited as file:line relative to
 * the EDKit.jl v0.5.0 source tree.  States are handled as 0-based packed
 * integers s = content - 1 (the reference stores 1-based contents, I[i]).
 *
 * Pinning (what checks each part independently of restating the formula):
 *   codec / index()        brute force over all base^L states at small L,
 *                          the intertwining identity O_full P = P H_sector
 *   enumerators            closed-form necklace/bracelet (Burnside) counts,
 *                          definition-based brute force at small L,
 *                          Gosper/Benes known answers from the reference tests
 *   scatter mul!           dense P^H O_full P, kron_embed, Hermiticity
 *   tridiagonal eigen      numpy.linalg.eigh, residual ||T q - lambda q||
 *   Lanczos / Krylov       dense eigvalsh / expm (scipy) on small sectors,
 *                          closed forms (4-site Heisenberg ring, XX chain)
 */
#ifndef EDK_ORACLE_H
#define EDK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_TENSOR = 0,
  ORC_PROJECTED = 1,
  ORC_TRANSLATIONAL = 2,
  ORC_TRANSLATION_PARITY = 3,
  ORC_ABELIAN = 4,
  ORC_PARITY = 5,            /* ParityBasis.jl */
  ORC_FLIP = 6,              /* FlipBasis.jl */
  ORC_PARITY_FLIP = 7,       /* ParityFlipBasis.jl */
  ORC_TRANSLATION_FLIP = 8   /* TranslationalFlipBasis.jl */
};

typedef struct orc_basis orc_basis;
typedef struct orc_op orc_op;
typedef struct orc_krylov orc_krylov;

/* Basis construction.
 *   kind   ORC_* above
 *   N      particle number as the reference's keyword (-1 = nothing)
 *   k, p   momentum / parity keywords; for ORC_ABELIAN has_k/has_p/has_z say
 *          which of basis(...)'s k/p/z keywords were given.
 *   ncustom, perms (ncustom*L, 1-based), ks, invs (ncustom*L bytes):
 *          basis(symmetries=[(perm,k[,inv])...]) generators.
 *   nthreads  worker threads for enumeration (result independent of it). */
orc_basis *orc_basis_build(int kind, int L, int base, int a, int N, int k,
                           int p, int z, int has_k, int has_p, int has_z,
                           int ncustom, const int32_t *perms,
                           const int32_t *ks, const uint8_t *invs,
                           int nthreads);
void orc_basis_free(orc_basis *b);
int64_t orc_basis_dim(const orc_basis *b);
/* I (1-based contents, exactly as the reference stores b.I) and R. */
void orc_basis_tables(const orc_basis *b, int64_t *I1, double *R);
/* index(b, dgt) for n packed 0-based states: 1-based pos and complex coef. */
/* schmidt(v, Ainds, b) (Schmidt.jl:186-330, AbelianBasis.jl:704-719): v is
 * D complex doubles, M the zeroed base^nA x base^(L-nA) complex matrix,
 * column-major; Ainds 1-based. */
void orc_schmidt(const orc_basis *b, const double *v, int nA, const int32_t *Ainds, double *M);
void orc_basis_index(const orc_basis *b, const uint64_t *states, int64_t n,
                     int64_t *pos1, double *coef_re, double *coef_im);
int orc_basis_order(const orc_basis *b);

/* Operator: flattened Vector{SparseMatrixCSC} + Vector{Vector{Int}}.
 *   site_off[nterms+1], sites[] 1-based
 *   colptr_off[nterms]: start of term t's colptr (length base^k_t + 1) in colptr[]
 *   colptr[]: 1-based positions into the term's own rowval/nzval slice
 *   nz_off[nterms+1], rowval[] 1-based, nzval[] complex interleaved (re, im) */
orc_op *orc_op_create(const orc_basis *b, int64_t nterms,
                      const int64_t *site_off, const int64_t *sites,
                      const int64_t *colptr_off, const int64_t *colptr,
                      const int64_t *nz_off, const int64_t *rowval,
                      const double *nzval);
void orc_op_free(orc_op *op);

/* mul!(target, opt, v): accumulates (target is not zeroed).  nthreads == 1 is
 * the serial mul!; nthreads > 1 is the threaded mul with per-thread private
 * full-length buffers summed at the end, added onto target.
 * target, v: complex interleaved, length D. */
void orc_mul(const orc_op *op, double *target, const double *v, int nthreads);
/* Block form mul!(Y, opt, X): X, Y row-major D x nb complex. */
void orc_mul_block(const orc_op *op, double *Y, const double *X, int64_t nb,
                   int nthreads);
/* Threaded mul over a subset of columns (bounded CPU-baseline samples). */
void orc_mul_cols(const orc_op *op, double *target, const double *v, const int64_t *cols,
                  int64_t ncols, int nthreads);
/* Column j (0-based) of the sector matrix: (1-based pos, complex value)
 * pairs, at most cap; returns the count (-1 if cap is too small). */
int64_t orc_column(const orc_op *op, int64_t j, int64_t *pos, double *val, int64_t cap);
/* Number of off-diagonal local-term nonzeros whose target lands in the
 * sector, summed over all rows (the NNZ_off of the byte model). */
int64_t orc_count_offdiag(const orc_op *op, int nthreads);

/* Symmetric tridiagonal eigensolver (implicit QL, Wilkinson shift).
 * alpha[m], beta[m-1]; outputs ascending lambda[m], Q column-major m x m. */
int orc_tridiag_eig(int m, const double *alpha, const double *beta,
                    double *lambda, double *Q);

/* Lanczos ground state (no counterpart in the reference; built from the
 * _lanczos_build! recurrence).  v0 complex interleaved, normalised inside.
 * Runs up to m steps, stops early when |beta_j Q[j,1]| <= tol (tol <= 0 runs
 * all m).  Returns steps taken; E0, resid; vec (optional, D complex). */
int orc_lanczos_ground(const orc_op *op, const double *v0, int m, double tol,
                       int nthreads, double *E0, double *resid, double *vec);

/* KrylovEvolutionCache restatement. */
orc_krylov *orc_krylov_create(const orc_op *op, const double *psi0,
                              double tol, int m_init, int m_max,
                              int extend_step, int nsample, int bisect_iters,
                              int extend_basis, int normalize_output,
                              int nthreads);
/* timeevolve!(cache, t) -> out (D complex).  Returns 0 on success, nonzero
 * with a message in orc_last_error() on the reference's error paths. */
int orc_krylov_evolve(orc_krylov *c, double t, double *out);
/* basis_builds, basis_extensions, restarts, matvecs, total_times_served,
 * max_dim_used, n_accepted */
void orc_krylov_diagnostics(const orc_krylov *c, int64_t *out7);
int orc_krylov_m(const orc_krylov *c);
void orc_krylov_free(orc_krylov *c);
const char *orc_last_error(void);

/* Known-answer helpers exposed for the reference's own unit tests. */
uint64_t orc_gosper_next(uint64_t s);
/* Benes network compiled from a 1-based perm and applied to a state. */
uint64_t orc_benes_apply(const int32_t *perm, int L, uint64_t state);
uint64_t orc_perm_apply_int(const int32_t *perm, int L, uint64_t state);
int orc_compute_period(const int32_t *perm, const uint8_t *inv, int L);

#ifdef __cplusplus
}
#endif
#endif
