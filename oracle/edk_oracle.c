/*
 * edk_oracle.c -- CPU restatement of EDKit.jl's matrix-free Operator path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * timed CPU baseline.  The product library never links or calls this file.
 *
 * Citations are file:line in the EDKit.jl v0.5.0 tree (src/...).  States are
 * 0-based packed integers s; the reference's 1-based content is s + 1.
 * Built with -ffp-contract=off so products and sums round exactly as written
 * (Julia does not contract a*b+c into an fma by default).
 */
#include "edk_oracle.h"

#include <complex.h>
#undef I
#define IU _Complex_I
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#define MAXL 64
#define MAXGEN 8

static __thread char g_err[256];
const char *orc_last_error(void) { return g_err; }
static void set_err(const char *m) {
  strncpy(g_err, m, sizeof(g_err) - 1);
  g_err[sizeof(g_err) - 1] = 0;
}

/* ------------------------------------------------------------------ */
/* Benes delta-swap network: AbelianBasis.jl:28-87                      */
/* ------------------------------------------------------------------ */
typedef struct {
  int n;
  uint64_t mask[MAXL];
  int shift[MAXL];
} benes_t;

/* compile_benes (AbelianBasis.jl:28-69): bit (L-i) moves to bit (L-perm[i]);
 * realised as an ordered list of transpositions found by placing, for each
 * destination bit in turn, the source bit that belongs there. */
static void benes_compile(const int32_t *perm, int L, benes_t *bn) {
  int bit_perm[MAXL], cur[MAXL], where[MAXL], inv[MAXL];
  for (int i = 1; i <= L; i++) bit_perm[L - i] = L - perm[i - 1];
  for (int i = 0; i < L; i++) {
    cur[i] = i;
    where[i] = i;
  }
  for (int s = 0; s < L; s++) inv[bit_perm[s]] = s;
  bn->n = 0;
  for (int dst = 0; dst < L; dst++) {
    int src = inv[dst];
    int cp = where[src];
    if (cp != dst) {
      int lo = dst < cp ? dst : cp, hi = dst < cp ? cp : dst;
      bn->mask[bn->n] = (uint64_t)1 << lo;
      bn->shift[bn->n] = hi - lo;
      bn->n++;
      int other = cur[dst];
      int t = cur[dst];
      cur[dst] = cur[cp];
      cur[cp] = t;
      where[src] = dst;
      where[other] = cp;
    }
  }
}

/* apply_benes (AbelianBasis.jl:80-87) */
static inline uint64_t benes_apply(const benes_t *bn, uint64_t s) {
  for (int i = 0; i < bn->n; i++) {
    uint64_t d = ((s >> bn->shift[i]) ^ s) & bn->mask[i];
    s = s ^ d ^ (d << bn->shift[i]);
  }
  return s;
}

/* apply_perm_int (AbelianBasis.jl:95-102) */
uint64_t orc_perm_apply_int(const int32_t *perm, int L, uint64_t s) {
  uint64_t r = 0;
  for (int i = 1; i <= L; i++) {
    uint64_t bit = (s >> (L - i)) & 1u;
    r |= bit << (L - perm[i - 1]);
  }
  return r;
}

uint64_t orc_benes_apply(const int32_t *perm, int L, uint64_t state) {
  benes_t bn;
  benes_compile(perm, L, &bn);
  return benes_apply(&bn, state);
}

/* _gosper_next (AbelianBasis.jl:456-463) */
uint64_t orc_gosper_next(uint64_t s) {
  uint64_t c = s & (~s + 1);
  uint64_t r = s + c;
  uint64_t diff = r ^ s;
  diff = (diff >> 2) / c;
  return r | diff;
}

/* _compute_period (AbelianBasis.jl:109-131) */
int orc_compute_period(const int32_t *perm, const uint8_t *inv, int L) {
  int pos[MAXL], np[MAXL];
  uint8_t fl[MAXL], nf[MAXL];
  for (int i = 0; i < L; i++) {
    pos[i] = i + 1;
    fl[i] = 0;
  }
  int lim = 2 * L * L > 100 ? 2 * L * L : 100;
  for (int p = 1; p <= lim; p++) {
    for (int i = 0; i < L; i++) {
      np[i] = perm[pos[i] - 1];
      nf[i] = fl[i] ^ (inv ? inv[pos[i] - 1] : 0);
    }
    int ident = 1;
    for (int i = 0; i < L; i++) {
      pos[i] = np[i];
      fl[i] = nf[i];
      if (pos[i] != i + 1 || fl[i]) ident = 0;
    }
    if (ident) return p;
  }
  return -1;
}

/* ------------------------------------------------------------------ */
/* Basis                                                              */
/* ------------------------------------------------------------------ */
struct orc_basis {
  int kind, L, base, A, len;
  int kp;  /* Translational: k' = mod(-k, len); TP: K = +-1 */
  int P;   /* TP parity; reflection or flip eigenvalue of the flip bases */
  int Z;   /* ParityFlipBasis: spin-flip eigenvalue */
  int Nnum; /* digit-sum constraint (-1 none) */
  int64_t D;
  int64_t *tab; /* 1-based contents (b.I) */
  double *R;
  cplx *C; /* Translational / TP phase table (len) */
  uint64_t pw[MAXL + 1];
  /* Abelian */
  int ngen, order;
  int gord[MAXGEN];
  int32_t gperm[MAXGEN][MAXL];
  uint8_t ginv[MAXGEN][MAXL];
  uint64_t ginvmask[MAXGEN];
  benes_t gb[MAXGEN];
  cplx *gc[MAXGEN];
  double *Cn; /* norm table indexed by stabiliser count (1-based) */
};

int64_t orc_basis_dim(const orc_basis *b) { return b->D; }
int orc_basis_order(const orc_basis *b) {
  switch (b->kind) {
    case ORC_TRANSLATIONAL: return b->L;        /* TranslationalBasis.jl:349 */
    case ORC_TRANSLATION_PARITY: return 2 * b->L; /* TranslationalParityBasis.jl:241 */
    case ORC_ABELIAN: return b->order;           /* AbelianBasis.jl:507 */
    case ORC_PARITY: return 2;                   /* ParityBasis.jl:128 */
    case ORC_FLIP: return 2;                     /* FlipBasis.jl:122 */
    case ORC_PARITY_FLIP: return 4;              /* ParityFlipBasis.jl:152 */
    case ORC_TRANSLATION_FLIP: return 2 * b->L;  /* TranslationalFlipBasis.jl:222 */
    default: return 1;                           /* AbstractBasis.jl:197 */
  }
}

/* _int_translate (AbstractBasis.jl:405-408) == circshift!(dgt, A) + index.
 * For base 2 the reference's own index() rotates with shifts
 * (TranslationalBasis.jl:318-325); the same value is computed here. */
static inline uint64_t translate(const orc_basis *b, uint64_t s) {
  if (b->base == 2) {
    int L = b->L, A = b->A;
    uint64_t mask = (L == 64) ? ~0ull : ((1ull << L) - 1);
    return ((s >> A) | (s << (L - A))) & mask;
  }
  uint64_t bs = b->pw[b->A];
  return (s % bs) * b->pw[b->L - b->A] + s / bs;
}

static inline uint64_t brev64(uint64_t v) {
  v = ((v >> 1) & 0x5555555555555555ull) | ((v & 0x5555555555555555ull) << 1);
  v = ((v >> 2) & 0x3333333333333333ull) | ((v & 0x3333333333333333ull) << 2);
  v = ((v >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((v & 0x0F0F0F0F0F0F0F0Full) << 4);
  v = ((v >> 8) & 0x00FF00FF00FF00FFull) | ((v & 0x00FF00FF00FF00FFull) << 8);
  v = ((v >> 16) & 0x0000FFFF0000FFFFull) | ((v & 0x0000FFFF0000FFFFull) << 16);
  return (v >> 32) | (v << 32);
}

/* _int_reverse (AbstractBasis.jl:417-425) == rindex - 1; for base 2 the
 * digit reversal is the reversal of the low L bits. */
static inline uint64_t reverse_digits(const orc_basis *b, uint64_t s) {
  if (b->base == 2) return brev64(s) >> (64 - b->L);
  uint64_t r = 0, v = s;
  for (int i = 0; i < b->L; i++) {
    r = r * b->base + v % b->base;
    v /= b->base;
  }
  return r;
}

/* _int_spinflip (AbstractBasis.jl:434) == MAX - i with MAX = base^L + 1 */
static inline uint64_t spinflip(const orc_basis *b, uint64_t s) { return b->pw[b->L] - 1 - s; }

static inline int digit_sum(const orc_basis *b, uint64_t s) {
  int t = 0;
  for (int i = 0; i < b->L; i++) {
    t += (int)(s % b->base);
    s /= b->base;
  }
  return t;
}

/* binary_search (ProjectedBasis.jl:71-81); returns 1-based pos or 0 */
static int64_t bsearch1(const int64_t *list, int64_t n, int64_t i) {
  int64_t l = 1, r = n, c = (l + r) / 2;
  if (n == 0) return 0;
  while (1) {
    int64_t t = list[c - 1];
    if (i < t) r = c - 1;
    else if (i > t) l = c + 1;
    else break;
    if (l > r) {
      c = 0;
      break;
    }
    c = (l + r) / 2;
  }
  return c;
}

/* Abelian group action on digits for base != 2 (AbelianBasis.jl:247-283) */
static uint64_t abelian_apply_digits(const orc_basis *b, int g, uint64_t s) {
  int L = b->L, d[MAXL], t[MAXL];
  uint64_t v = s;
  for (int i = L - 1; i >= 0; i--) {
    d[i] = (int)(v % b->base);
    v /= b->base;
  }
  for (int i = 0; i < L; i++) t[i] = d[i];
  for (int i = 0; i < L; i++) d[b->gperm[g][i] - 1] = t[i]; /* apply_perm! */
  for (int i = 0; i < L; i++)
    if (b->ginv[g][i]) d[i] = b->base - 1 - d[i]; /* apply_inv! */
  uint64_t r = 0;
  for (int i = 0; i < L; i++) r = r * b->base + d[i];
  return r;
}

/* _apply_generator_int (AbelianBasis.jl:380-389) / digit equivalent */
static inline uint64_t abelian_gen(const orc_basis *b, int g, uint64_t s) {
  if (b->base == 2) return benes_apply(&b->gb[g], s) ^ b->ginvmask[g];
  return abelian_apply_digits(b, g, s);
}

/* One odometer step of the group iteration (AbelianBasis.jl:286-299,
 * 406-410): apply generator 1; on wrap-around carry into generator 2, ... */
static inline uint64_t odometer_step(const orc_basis *b, int *os, uint64_t cur) {
  for (int i = 0; i < b->ngen; i++) {
    cur = abelian_gen(b, i, cur);
    if (++os[i] > b->gord[i]) os[i] = 1;
    else break;
  }
  return cur;
}

/* phase (AbelianBasis.jl:225-227): product over generators, left to right */
static inline cplx abelian_phase(const orc_basis *b, const int *os) {
  cplx ph = b->gc[0][os[0] - 1];
  for (int i = 1; i < b->ngen; i++) ph = ph * b->gc[i][os[i] - 1];
  return ph;
}

/* Judges: return 1 and the norm when the candidate is a representative. */

/* TranslationJudge (TranslationalBasis.jl:94-114) */
static int judge_trans(const orc_basis *b, uint64_t s, double *nrm) {
  uint64_t st = s;
  for (int n = 1; n <= b->len - 1; n++) {
    st = translate(b, st);
    if (st < s) return 0;
    if (st == s) {
      if (((int64_t)n * b->kp) % b->len != 0) return 0; /* check_momentum :80 */
      *nrm = (double)b->len / sqrt((double)n);         /* norm table :250 */
      return 1;
    }
  }
  *nrm = (double)b->len / sqrt((double)b->len);
  return 1;
}

/* TranslationParityJudge (TranslationalParityBasis.jl:68-112) */
static int judge_tp(const orc_basis *b, uint64_t s, double *nrm) {
  int R = b->len, Qp = 0;
  uint64_t In = reverse_digits(b, s);
  if (In < s) return 0;
  if (In == s) {
    Qp = 1;
    if (b->P != 1) return 0;
  }
  uint64_t st = s;
  for (int n = 1; n <= b->len - 1; n++) {
    st = translate(b, st);
    uint64_t In1 = st, In2 = reverse_digits(b, st);
    if (In1 < s) return 0;
    if (In2 < s) return 0;
    if (In2 == s) {
      int Kn = (b->kp == 1 || (n % 2 == 0)) ? 1 : -1;
      if (b->P * Kn != 1) return 0;
      Qp = 1;
    }
    if (In1 == s) {
      if (!(b->kp == 1 || n % 2 == 0)) return 0;
      R = n;
      break;
    }
  }
  /* N2 = 2len/sqrt(i), N1 = N2/sqrt(2)   (TranslationalParityBasis.jl:158-159) */
  double n2 = (double)(2 * b->len) / sqrt((double)R);
  *nrm = Qp ? n2 : n2 / sqrt(2.0);
  return 1;
}

/* check_min_int (AbelianBasis.jl:399-418) and check_min (:322-338) */
static int judge_abelian(const orc_basis *b, uint64_t s, double *nrm) {
  int os[MAXGEN];
  for (int i = 0; i < b->ngen; i++) os[i] = 1;
  uint64_t cur = s;
  int N = 1;
  for (int st = 2; st <= b->order; st++) {
    cur = odometer_step(b, os, cur);
    if (cur < s) return 0;
    if (cur == s) {
      if (!(cabs(abelian_phase(b, os) - 1.0) < 1e-14)) return 0;
      N++;
    }
  }
  *nrm = b->Cn[N - 1];
  return 1;
}

/* ParityJudge (ParityBasis.jl:44-57) and FlipJudge (FlipBasis.jl:42-55) */
static int judge_pair(const orc_basis *b, uint64_t s, double *nrm) {
  uint64_t In = (b->kind == ORC_PARITY) ? reverse_digits(b, s) : spinflip(b, s);
  if (In < s) return 0;
  if (In == s) {
    if (b->P != 1) return 0;
    *nrm = 2.0;
  } else {
    *nrm = sqrt(2.0);
  }
  return 1;
}

/* double_parity_check + ParityFlipJudge (ParityFlipBasis.jl:47-65, 73-79) */
static int judge_parityflip(const orc_basis *b, uint64_t s, double *nrm) {
  static const double Cn[4] = {2.0, 2.0 * 1.4142135623730951, 0.0, 4.0}; /* :96 */
  uint64_t Ip = reverse_digits(b, s);
  if (Ip < s) return 0;
  uint64_t Iz = spinflip(b, s);
  if (Iz < s) return 0;
  uint64_t Ipz = spinflip(b, Ip);
  if (Ipz < s) return 0;
  int N = 1;
  if (Ip == s) {
    if (b->P != 1) return 0;
    N++;
  }
  if (Iz == s) {
    if (b->Z != 1) return 0;
    N++;
  }
  if (Ipz == s) {
    if (b->P * b->Z != 1) return 0;
    N++;
  }
  *nrm = Cn[N - 1];
  return 1;
}

/* TranslationFlipJudge (TranslationalFlipBasis.jl:50-94) */
static int judge_transflip(const orc_basis *b, uint64_t s, double *nrm) {
  int R = b->len, Qp = 0;
  uint64_t In = spinflip(b, s);
  if (In < s) return 0;
  if (In == s) {
    Qp = 1;
    if (b->P != 1) return 0;
  }
  uint64_t st = s;
  for (int n = 1; n <= b->len - 1; n++) {
    st = translate(b, st); /* circshift!(dgt, A) */
    uint64_t In1 = st, In2 = spinflip(b, st);
    if (In1 < s) return 0;
    if (In2 < s) return 0;
    if (In2 == s) {
      /* isequal(isone(P), iseven(div(n * K, L / 2))) */
      int ev = ((n * b->kp) / (b->len / 2)) % 2 == 0;
      if ((b->P == 1) != ev) return 0;
      Qp = 1;
    }
    if (In1 == s) {
      if ((n * b->kp) % b->len != 0) return 0;
      R = n;
      break;
    }
  }
  double n2 = (double)(2 * b->len) / sqrt((double)R); /* N2, N1 = N2/sqrt(2) (:140-141) */
  *nrm = Qp ? n2 : n2 / sqrt(2.0);
  return 1;
}

static int judge(const orc_basis *b, uint64_t s, double *nrm) {
  switch (b->kind) {
    case ORC_PROJECTED: *nrm = 1.0; return 1;
    case ORC_TRANSLATIONAL: return judge_trans(b, s, nrm);
    case ORC_TRANSLATION_PARITY: return judge_tp(b, s, nrm);
    case ORC_ABELIAN: return judge_abelian(b, s, nrm);
    case ORC_PARITY:
    case ORC_FLIP: return judge_pair(b, s, nrm);
    case ORC_PARITY_FLIP: return judge_parityflip(b, s, nrm);
    case ORC_TRANSLATION_FLIP: return judge_transflip(b, s, nrm);
  }
  return 0;
}

/* ---- enumeration (selectindexnorm / _abelian_select*), threaded ---- */
static uint64_t binom_tab[MAXL + 1][MAXL + 1];
static void binom_init(void) {
  if (binom_tab[0][0]) return;
  for (int n = 0; n <= MAXL; n++) {
    binom_tab[n][0] = 1;
    for (int k = 1; k <= n; k++)
      binom_tab[n][k] = binom_tab[n - 1][k - 1] + (k <= n - 1 ? binom_tab[n - 1][k] : 0);
  }
}
/* r-th (0-based) smallest L-bit integer with n set bits */
static uint64_t unrank(uint64_t r, int L, int n) {
  uint64_t s = 0;
  for (int pos = L - 1; pos >= 0 && n > 0; pos--) {
    uint64_t c = binom_tab[pos][n]; /* states with this bit clear */
    if (r >= c) {
      s |= (uint64_t)1 << pos;
      r -= c;
      n--;
    }
  }
  return s;
}

typedef struct {
  const orc_basis *b;
  int gosper; /* 1: popcount enumeration */
  int pop;
  uint64_t lo, hi; /* candidate rank range (gosper) or value range */
  int64_t n, cap;
  int64_t *tab;
  double *R;
} enum_job;

static void job_push(enum_job *j, uint64_t s, double r) {
  if (j->n == j->cap) {
    j->cap = j->cap ? 2 * j->cap : 1024;
    j->tab = (int64_t *)realloc(j->tab, j->cap * sizeof(int64_t));
    j->R = (double *)realloc(j->R, j->cap * sizeof(double));
  }
  j->tab[j->n] = (int64_t)s + 1;
  j->R[j->n] = r;
  j->n++;
}

static void *enum_worker(void *arg) {
  enum_job *j = (enum_job *)arg;
  const orc_basis *b = j->b;
  double r;
  if (j->gosper) {
    if (j->lo >= j->hi) return NULL;
    uint64_t s = unrank(j->lo, b->L, j->pop);
    for (uint64_t q = j->lo; q < j->hi; q++) {
      if (judge(b, s, &r)) job_push(j, s, r);
      if (q + 1 < j->hi) s = orc_gosper_next(s);
    }
  } else {
    for (uint64_t s = j->lo; s < j->hi; s++) {
      if (b->Nnum >= 0 && digit_sum(b, s) != b->Nnum) continue;
      if (judge(b, s, &r)) job_push(j, s, r);
    }
  }
  return NULL;
}

static void enumerate(orc_basis *b, int gosper, int pop, int nthreads) {
  binom_init();
  uint64_t total = gosper ? binom_tab[b->L][pop] : b->pw[b->L];
  if (gosper && pop == 0) total = 1;
  if (nthreads < 1) nthreads = 1;
  if ((uint64_t)nthreads > total) nthreads = (int)(total ? total : 1);
  enum_job *jobs = (enum_job *)calloc(nthreads, sizeof(enum_job));
  pthread_t *th = (pthread_t *)calloc(nthreads, sizeof(pthread_t));
  uint64_t each = total / nthreads, left = total % nthreads, start = 0;
  for (int t = 0; t < nthreads; t++) {
    uint64_t cnt = each + ((uint64_t)t < left ? 1 : 0);
    jobs[t].b = b;
    jobs[t].gosper = gosper;
    jobs[t].pop = pop;
    jobs[t].lo = start;
    jobs[t].hi = start + cnt;
    start += cnt;
  }
  for (int t = 1; t < nthreads; t++) pthread_create(&th[t], NULL, enum_worker, &jobs[t]);
  enum_worker(&jobs[0]);
  for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
  int64_t D = 0;
  for (int t = 0; t < nthreads; t++) D += jobs[t].n;
  b->D = D;
  b->tab = (int64_t *)malloc((D ? D : 1) * sizeof(int64_t));
  b->R = (double *)malloc((D ? D : 1) * sizeof(double));
  int64_t o = 0;
  for (int t = 0; t < nthreads; t++) { /* vcat keeps chunk order: sorted */
    if (jobs[t].n) {
      memcpy(b->tab + o, jobs[t].tab, jobs[t].n * sizeof(int64_t));
      memcpy(b->R + o, jobs[t].R, jobs[t].n * sizeof(double));
    }
    o += jobs[t].n;
    free(jobs[t].tab);
    free(jobs[t].R);
  }
  free(jobs);
  free(th);
}

static void add_gen(orc_basis *b, int order, int k, const int32_t *perm, const uint8_t *inv) {
  /* AbelianOperator(order, k, perm; inv) (AbelianBasis.jl:167-193) */
  int g = b->ngen++;
  b->gord[g] = order;
  for (int i = 0; i < b->L; i++) {
    b->gperm[g][i] = perm[i];
    b->ginv[g][i] = inv ? inv[i] : 0;
  }
  b->gc[g] = (cplx *)malloc(order * sizeof(cplx));
  for (int j = 0; j < order; j++) {
    if (k == 0) b->gc[g][j] = 1.0;
    else if (2 * k == order) b->gc[g][j] = (j % 2 == 0) ? 1.0 : -1.0;
    else {
      double th = ((2.0 * M_PI * k) / order) * j;
      b->gc[g][j] = cos(th) + IU * sin(th);
    }
  }
  benes_compile(perm, b->L, &b->gb[g]);
  uint64_t m = 0;
  for (int i = 1; i <= b->L; i++)
    if (b->ginv[g][i - 1]) m |= (uint64_t)1 << (b->L - i);
  b->ginvmask[g] = m;
}

orc_basis *orc_basis_build(int kind, int L, int base, int a, int N, int k,
                           int p, int z, int has_k, int has_p, int has_z,
                           int ncustom, const int32_t *perms,
                           const int32_t *ks, const uint8_t *invs,
                           int nthreads) {
  if (L < 1 || L > 63 || base < 2) {
    set_err("unsupported L/base");
    return NULL;
  }
  orc_basis *b = (orc_basis *)calloc(1, sizeof(orc_basis));
  b->kind = kind;
  b->L = L;
  b->base = base;
  b->A = a > 0 ? a : 1;
  b->Nnum = -1;
  b->pw[0] = 1;
  for (int i = 1; i <= L; i++) b->pw[i] = b->pw[i - 1] * (uint64_t)base;
  if (L % b->A) {
    set_err("Length of unit-cell incompatible with L");
    free(b);
    return NULL;
  }
  b->len = L / b->A;
  int num = (N >= 0) ? L * (base - 1) - N : -1; /* sum(dgt) == L(B-1) - N */
  if (kind == ORC_ABELIAN) {
    /* basis(...) front door (AbelianBasis.jl:771-847) */
    int32_t perm[MAXL];
    if (has_k) {
      for (int i = 1; i <= L; i++) perm[i - 1] = ((i - b->A - 1) % L + L) % L + 1;
      add_gen(b, L / b->A, k, perm, NULL);
    }
    if (has_p) {
      for (int i = 1; i <= L; i++) perm[i - 1] = L - i + 1;
      add_gen(b, 2, p == -1 ? 1 : 0, perm, NULL);
    }
    if (has_z) {
      uint8_t inv[MAXL];
      for (int i = 1; i <= L; i++) {
        perm[i - 1] = i;
        inv[i - 1] = 1;
      }
      add_gen(b, 2, z == -1 ? 1 : 0, perm, inv);
    }
    for (int c = 0; c < ncustom; c++) {
      int ord = orc_compute_period(perms + c * L, invs ? invs + c * L : NULL, L);
      if (ord < 1) {
        set_err("Period computation failed");
        orc_basis_free(b);
        return NULL;
      }
      add_gen(b, ord, ks[c], perms + c * L, invs ? invs + c * L : NULL);
    }
    if (b->ngen == 0) {
      /* no generators: TensorBasis or ProjectedBasis */
      kind = b->kind = (N < 0) ? ORC_TENSOR : ORC_PROJECTED;
    } else {
      b->order = 1;
      for (int g = 0; g < b->ngen; g++) b->order *= b->gord[g];
      b->Cn = (double *)calloc(b->order, sizeof(double));
      for (int i = 1; i <= b->order; i++) /* AbelianBasis.jl:524-527 */
        if (b->order % i == 0) b->Cn[i - 1] = sqrt((double)(b->order * i));
      if (base == 2 && N >= 0 && N <= L) {
        enumerate(b, 1, N, nthreads); /* Gosper: N set bits, f dropped (:530-540) */
      } else {
        b->Nnum = num;
        enumerate(b, 0, 0, nthreads);
      }
      return b;
    }
  }
  if (kind == ORC_TENSOR) {
    b->D = (int64_t)b->pw[L];
    return b;
  }
  if (kind == ORC_TRANSLATIONAL) {
    b->kp = ((-k) % b->len + b->len) % b->len; /* TranslationalBasis.jl:247 */
    b->C = (cplx *)malloc(b->len * sizeof(cplx));
    for (int i = 0; i < b->len; i++) { /* phase_factor :277-285 */
      if (b->kp == 0) b->C[i] = 1.0;
      else if (2 * b->kp == b->len) b->C[i] = (i % 2 == 0) ? 1.0 : -1.0;
      else {
        double th = ((-2.0 * M_PI / b->len) * b->kp) * i;
        b->C[i] = cos(th) + IU * sin(th);
      }
    }
  } else if (kind == ORC_TRANSLATION_PARITY) {
    int km = ((k % b->len) + b->len) % b->len;
    if (km == 0) b->kp = 1; /* TranslationalParityBasis.jl:147-153 */
    else if ((2 * km) % b->len == 0) b->kp = -1;
    else {
      set_err("Momentum incompatible with parity.");
      free(b);
      return NULL;
    }
    if (p != 1 && p != -1) {
      set_err("Invalid parity");
      free(b);
      return NULL;
    }
    b->P = p;
    b->C = (cplx *)malloc(b->len * sizeof(cplx));
    for (int i = 0; i < b->len; i++) b->C[i] = (b->kp == 1 || i % 2 == 0) ? 1.0 : -1.0;
  } else if (kind == ORC_PARITY || kind == ORC_FLIP || kind == ORC_PARITY_FLIP ||
             kind == ORC_TRANSLATION_FLIP) {
    if ((p != 1 && p != -1) || (kind == ORC_PARITY_FLIP && z != 1 && z != -1)) {
      set_err("Invalid parity");
      free(b);
      return NULL;
    }
    if (kind != ORC_PARITY && N >= 0 && 2 * N != L * (base - 1)) {
      set_err("N not compatible."); /* FlipBasis.jl:77 and the like */
      free(b);
      return NULL;
    }
    b->P = p;
    b->Z = z;
    if (kind == ORC_TRANSLATION_FLIP) {
      b->kp = ((-k) % b->len + b->len) % b->len; /* TranslationalFlipBasis.jl:136 */
      b->C = (cplx *)malloc(b->len * sizeof(cplx));
      for (int i = 0; i < b->len; i++) { /* phase_factor (TranslationalBasis.jl:277-285) */
        if (b->kp == 0) b->C[i] = 1.0;
        else if (2 * b->kp == b->len) b->C[i] = (i % 2 == 0) ? 1.0 : -1.0;
        else {
          double th = ((-2.0 * M_PI / b->len) * b->kp) * i;
          b->C[i] = cos(th) + IU * sin(th);
        }
      }
    }
  } else if (kind != ORC_PROJECTED) {
    set_err("unknown basis kind");
    free(b);
    return NULL;
  }
  if (kind == ORC_PROJECTED && N < 0) {
    set_err("ProjectedBasis needs N (predicate f is not supported)");
    free(b);
    return NULL;
  }
  if (num >= 0 && base == 2) {
    enumerate(b, 1, num, nthreads); /* same set and order as the full scan */
  } else {
    b->Nnum = num;
    enumerate(b, 0, 0, nthreads);
  }
  return b;
}

void orc_basis_free(orc_basis *b) {
  if (!b) return;
  free(b->tab);
  free(b->R);
  free(b->C);
  free(b->Cn);
  for (int g = 0; g < b->ngen; g++) free(b->gc[g]);
  free(b);
}

void orc_basis_tables(const orc_basis *b, int64_t *I1, double *R) {
  if (b->kind == ORC_TENSOR) {
    for (int64_t i = 0; i < b->D; i++) {
      if (I1) I1[i] = i + 1;
      if (R) R[i] = 1.0;
    }
    return;
  }
  if (I1) memcpy(I1, b->tab, b->D * sizeof(int64_t));
  if (R) memcpy(R, b->R, b->D * sizeof(double));
}

/* index(b, dgt) -> (coef, pos); pos 1-based; (0, 1) when absent. */
static inline void basis_index(const orc_basis *b, uint64_t s, cplx *coef, int64_t *pos) {
  switch (b->kind) {
    case ORC_TENSOR: /* AbstractBasis.jl:246 */
      *coef = 1.0;
      *pos = (int64_t)s + 1;
      return;
    case ORC_PROJECTED: { /* ProjectedBasis.jl:52-57 */
      int64_t i = bsearch1(b->tab, b->D, (int64_t)s + 1);
      if (i > 0) {
        *coef = 1.0;
        *pos = i;
      } else {
        *coef = 0.0;
        *pos = 1;
      }
      return;
    }
    case ORC_TRANSLATIONAL: { /* TranslationalBasis.jl:310-341 */
      uint64_t I0 = s + 1, Im = I0, st = s;
      int M = 0;
      for (int i = 1; i <= b->len - 1; i++) {
        st = translate(b, st);
        uint64_t In = st + 1;
        if (In == I0) break;
        if (In < Im) {
          Im = In;
          M = i;
        }
      }
      int64_t i = bsearch1(b->tab, b->D, (int64_t)Im);
      if (i == 0) {
        *coef = 0.0;
        *pos = 1;
        return;
      }
      *coef = b->C[M] * b->R[i - 1];
      *pos = i;
      return;
    }
    case ORC_TRANSLATION_PARITY: { /* TranslationalParityBasis.jl:192-234 */
      uint64_t Ia0 = s + 1, st = s;
      uint64_t Ib0 = reverse_digits(b, s) + 1;
      uint64_t Iam = Ia0, Ibm = Ib0;
      int R = 0, M = 0;
      for (int i = 1; i <= b->len - 1; i++) {
        st = translate(b, st);
        uint64_t Ia1 = st + 1;
        if (Ia1 == Ia0) break;
        else if (Ia1 < Iam) {
          Iam = Ia1;
          R = i;
        }
        uint64_t Ib1 = reverse_digits(b, st) + 1;
        if (Ib1 < Ibm) {
          Ibm = Ib1;
          M = i;
        }
      }
      int64_t i;
      cplx n;
      if (Ibm < Iam) {
        i = bsearch1(b->tab, b->D, (int64_t)Ibm);
        n = (double)b->P * b->C[M];
      } else {
        i = bsearch1(b->tab, b->D, (int64_t)Iam);
        n = b->C[R];
      }
      if (i == 0) {
        *coef = 0.0;
        *pos = 1;
        return;
      }
      *coef = n * b->R[i - 1];
      *pos = i;
      return;
    }
    case ORC_PARITY:   /* ParityBasis.jl:110-121 */
    case ORC_FLIP: {   /* FlipBasis.jl:105-115 */
      uint64_t Ia = s + 1;
      uint64_t Ib = ((b->kind == ORC_PARITY) ? reverse_digits(b, s) : spinflip(b, s)) + 1;
      int64_t i;
      double n;
      if (Ib < Ia) {
        i = bsearch1(b->tab, b->D, (int64_t)Ib);
        n = (double)b->P;
      } else {
        i = bsearch1(b->tab, b->D, (int64_t)Ia);
        n = 1.0;
      }
      if (i == 0) {
        *coef = 0.0;
        *pos = 1;
        return;
      }
      *coef = n * b->R[i - 1];
      *pos = i;
      return;
    }
    case ORC_PARITY_FLIP: { /* ParityFlipBasis.jl:122-145 */
      uint64_t I0 = s + 1, Ip = reverse_digits(b, s) + 1, Iz = spinflip(b, s) + 1;
      uint64_t Ipz = spinflip(b, Ip - 1) + 1;
      uint64_t Ir = I0;
      if (Iz < Ir) Ir = Iz;
      if (Ip < Ir) Ir = Ip;
      if (Ipz < Ir) Ir = Ipz;
      int64_t i = bsearch1(b->tab, b->D, (int64_t)Ir);
      if (i == 0) {
        *coef = 0.0;
        *pos = 1;
        return;
      }
      double n;
      if (Ir == I0) n = 1.0;
      else if (Ir == Ip) n = (double)b->P;
      else if (Ir == Iz) n = (double)b->Z;
      else n = (double)(b->P * b->Z);
      *coef = n * b->R[i - 1];
      *pos = i;
      return;
    }
    case ORC_TRANSLATION_FLIP: { /* TranslationalFlipBasis.jl:172-215 */
      uint64_t Ia0 = s + 1, st = s;
      uint64_t Ib0 = spinflip(b, s) + 1;
      uint64_t Iam = Ia0, Ibm = Ib0;
      int R = 0, M = 0;
      for (int i = 1; i <= b->len - 1; i++) {
        st = translate(b, st);
        uint64_t Ia1 = st + 1;
        if (Ia1 == Ia0) break;
        else if (Ia1 < Iam) {
          Iam = Ia1;
          R = i;
        }
        uint64_t Ib1 = spinflip(b, st) + 1;
        if (Ib1 < Ibm) {
          Ibm = Ib1;
          M = i;
        }
      }
      int64_t i;
      cplx n;
      if (Ibm < Iam) {
        i = bsearch1(b->tab, b->D, (int64_t)Ibm);
        n = (double)b->P * b->C[M];
      } else {
        i = bsearch1(b->tab, b->D, (int64_t)Iam);
        n = b->C[R];
      }
      if (i == 0) {
        *coef = 0.0;
        *pos = 1;
        return;
      }
      *coef = n * b->R[i - 1];
      *pos = i;
      return;
    }
    case ORC_ABELIAN: { /* AbelianBasis.jl:669-693 + shift_canonical_int :428-445 */
      int os[MAXGEN], ms[MAXGEN];
      for (int g = 0; g < b->ngen; g++) os[g] = ms[g] = 1;
      uint64_t best = s, cur = s;
      for (int st = 1; st <= b->order; st++) {
        cur = odometer_step(b, os, cur);
        if (cur < best) {
          best = cur;
          for (int g = 0; g < b->ngen; g++) ms[g] = os[g];
        }
      }
      int64_t i = bsearch1(b->tab, b->D, (int64_t)best + 1);
      if (i == 0) {
        *coef = 0.0;
        *pos = 1;
        return;
      }
      *coef = abelian_phase(b, ms) * b->R[i - 1];
      *pos = i;
      return;
    }
  }
}

/* ---- schmidt (Schmidt.jl:186-330, AbelianBasis.jl:704-719) ---- */
/* addto! (Schmidt.jl:65-72): M[index(A digits), index(B digits)] += val, the
 * digits of site q (1-based) being the q-th most significant of the content */
static void schmidt_add(const orc_basis *b, const int *Ainds, int nA, const int *Binds, int nB, int64_t dA,
                        cplx *M, uint64_t s, cplx val) {
  int d[MAXL];
  uint64_t v = s;
  for (int i = b->L - 1; i >= 0; i--) {
    d[i] = (int)(v % b->base);
    v /= b->base;
  }
  int64_t ia = 0, ib = 0;
  for (int q = 0; q < nA; q++) ia = ia * b->base + d[Ainds[q] - 1];
  for (int q = 0; q < nB; q++) ib = ib * b->base + d[Binds[q] - 1];
  M[ia + dA * ib] += val;
}

/* M: base^nA x base^(L-nA), column-major, zeroed by the caller */
void orc_schmidt(const orc_basis *b, const double *v, int nA, const int32_t *Ainds, double *Mout) {
  cplx *M = (cplx *)Mout;
  const cplx *x = (const cplx *)v;
  int A[MAXL], B[MAXL], nB = 0, in[MAXL + 1] = {0};
  for (int q = 0; q < nA; q++) {
    A[q] = Ainds[q];
    in[Ainds[q]] = 1;
  }
  for (int q = 1; q <= b->L; q++) /* Binds (Schmidt.jl:46-52) */
    if (!in[q]) B[nB++] = q;
  int64_t dA = 1;
  for (int q = 0; q < nA; q++) dA *= b->base;
  if (b->kind == ORC_TENSOR || b->kind == ORC_PROJECTED) { /* Schmidt.jl:186-195 */
    for (int64_t i = 0; i < b->D; i++) {
      uint64_t s = b->kind == ORC_TENSOR ? (uint64_t)i : (uint64_t)(b->tab[i] - 1);
      schmidt_add(b, A, nA, B, nB, dA, M, s, x[i]);
    }
    return;
  }
  /* b.C[2] */
  cplx step = 1.0;
  if ((b->kind == ORC_TRANSLATIONAL || b->kind == ORC_TRANSLATION_PARITY || b->kind == ORC_TRANSLATION_FLIP) &&
      b->len > 1)
    step = b->C[1];
  for (int64_t i = 0; i < b->D; i++) {
    uint64_t s = (uint64_t)(b->tab[i] - 1);
    cplx val = x[i] / b->R[i];
    switch (b->kind) {
      case ORC_TRANSLATIONAL:       /* Schmidt.jl:202-216 */
      case ORC_TRANSLATION_PARITY:  /* parity_schmidt :249-273 */
      case ORC_TRANSLATION_FLIP:
        for (int j = 0; j < b->len; j++) {
          schmidt_add(b, A, nA, B, nB, dA, M, s, val);
          s = translate(b, s);
          val *= step;
        }
        if (b->kind == ORC_TRANSLATIONAL) break;
        s = b->kind == ORC_TRANSLATION_PARITY ? reverse_digits(b, s) : spinflip(b, s);
        val *= (double)b->P;
        for (int j = 0; j < b->len; j++) {
          schmidt_add(b, A, nA, B, nB, dA, M, s, val);
          s = translate(b, s);
          val *= step;
        }
        break;
      case ORC_PARITY: /* Schmidt.jl:296-309 */
      case ORC_FLIP:   /* Schmidt.jl:279-292 */
        schmidt_add(b, A, nA, B, nB, dA, M, s, val);
        s = b->kind == ORC_PARITY ? reverse_digits(b, s) : spinflip(b, s);
        schmidt_add(b, A, nA, B, nB, dA, M, s, (double)b->P * val);
        break;
      case ORC_PARITY_FLIP: /* Schmidt.jl:313-330 */
        schmidt_add(b, A, nA, B, nB, dA, M, s, val);
        s = reverse_digits(b, s);
        val *= (double)b->P;
        schmidt_add(b, A, nA, B, nB, dA, M, s, val);
        s = spinflip(b, s);
        val *= (double)b->Z;
        schmidt_add(b, A, nA, B, nB, dA, M, s, val);
        s = reverse_digits(b, s);
        val *= (double)b->P;
        schmidt_add(b, A, nA, B, nB, dA, M, s, val);
        break;
      case ORC_ABELIAN: { /* AbelianBasis.jl:704-719 */
        int os[MAXGEN];
        for (int g = 0; g < b->ngen; g++) os[g] = 1;
        schmidt_add(b, A, nA, B, nB, dA, M, s, val);
        for (int st = 2; st <= b->order; st++) {
          s = odometer_step(b, os, s);
          schmidt_add(b, A, nA, B, nB, dA, M, s, abelian_phase(b, os) * val);
        }
        break;
      }
    }
  }
}

void orc_basis_index(const orc_basis *b, const uint64_t *states, int64_t n,
                     int64_t *pos1, double *cre, double *cim) {
  for (int64_t q = 0; q < n; q++) {
    cplx c;
    int64_t p;
    basis_index(b, states[q], &c, &p);
    pos1[q] = p;
    cre[q] = creal(c);
    cim[q] = cimag(c);
  }
}

/* ------------------------------------------------------------------ */
/* Operator + scatter matvec: Operator.jl:24-78, 309-465                */
/* ------------------------------------------------------------------ */
struct orc_op {
  const orc_basis *b;
  int64_t nterms;
  int *nsite;
  int (*sites)[MAXL]; /* 1-based */
  int64_t **colptr;   /* 0-based, per term */
  int64_t **rowval;   /* 0-based */
  cplx **nzval;
};

orc_op *orc_op_create(const orc_basis *b, int64_t nterms,
                      const int64_t *site_off, const int64_t *sites,
                      const int64_t *colptr_off, const int64_t *colptr,
                      const int64_t *nz_off, const int64_t *rowval,
                      const double *nzval) {
  orc_op *op = (orc_op *)calloc(1, sizeof(orc_op));
  op->b = b;
  op->nterms = nterms;
  op->nsite = (int *)calloc(nterms ? nterms : 1, sizeof(int));
  op->sites = calloc(nterms ? nterms : 1, sizeof(*op->sites));
  op->colptr = (int64_t **)calloc(nterms ? nterms : 1, sizeof(int64_t *));
  op->rowval = (int64_t **)calloc(nterms ? nterms : 1, sizeof(int64_t *));
  op->nzval = (cplx **)calloc(nterms ? nterms : 1, sizeof(cplx *));
  for (int64_t t = 0; t < nterms; t++) {
    int ns = (int)(site_off[t + 1] - site_off[t]);
    op->nsite[t] = ns;
    for (int q = 0; q < ns; q++) op->sites[t][q] = (int)sites[site_off[t] + q];
    uint64_t n = 1;
    for (int q = 0; q < ns; q++) n *= (uint64_t)b->base;
    op->colptr[t] = (int64_t *)malloc((n + 1) * sizeof(int64_t));
    for (uint64_t c = 0; c <= n; c++) op->colptr[t][c] = colptr[colptr_off[t] + c] - 1;
    int64_t nnz = nz_off[t + 1] - nz_off[t];
    op->rowval[t] = (int64_t *)malloc((nnz ? nnz : 1) * sizeof(int64_t));
    op->nzval[t] = (cplx *)malloc((nnz ? nnz : 1) * sizeof(cplx));
    for (int64_t q = 0; q < nnz; q++) {
      op->rowval[t][q] = rowval[nz_off[t] + q] - 1;
      op->nzval[t][q] = nzval[2 * (nz_off[t] + q)] + IU * nzval[2 * (nz_off[t] + q) + 1];
    }
  }
  return op;
}

void orc_op_free(orc_op *op) {
  if (!op) return;
  for (int64_t t = 0; t < op->nterms; t++) {
    free(op->colptr[t]);
    free(op->rowval[t]);
    free(op->nzval[t]);
  }
  free(op->nsite);
  free(op->sites);
  free(op->colptr);
  free(op->rowval);
  free(op->nzval);
  free(op);
}

static inline uint64_t content0(const orc_basis *b, int64_t j) {
  return b->kind == ORC_TENSOR ? (uint64_t)j : (uint64_t)(b->tab[j] - 1); /* :95, :240 */
}
static inline double basis_norm(const orc_basis *b, int64_t j) {
  return (b->kind == ORC_TENSOR || b->kind == ORC_PROJECTED) ? 1.0 : b->R[j]; /* :106-107 */
}

/* colmn!(target, opt, j, dgt, coeff) (Operator.jl:458-465) over the local
 * terms, each colmn!(target, M, I, b, dgt, C) (:426-439).  nb columns of a
 * row-major block are handled together like _accumulate! (:399-407). */
static void colmn(const orc_op *op, cplx *target, int64_t j, const cplx *coeff, int64_t nb) {
  const orc_basis *b = op->b;
  uint64_t s = content0(b, j);
  double r = basis_norm(b, j);
  cplx C[64];
  for (int64_t q = 0; q < nb; q++) C[q] = (r == 1.0) ? coeff[q] : coeff[q] / r;
  for (int64_t t = 0; t < op->nterms; t++) {
    int ns = op->nsite[t];
    const int *S = op->sites[t];
    /* jl = index(dgt, I) (AbstractBasis.jl:318-331), 0-based here */
    uint64_t jl = 0;
    for (int q = 0; q < ns; q++) {
      uint64_t w = b->pw[b->L - S[q]];
      uint64_t d = (s / w) % b->base;
      jl = jl * b->base + d;
    }
    /* Repeated sites in one term keep the reference's last-write order. */
    for (int64_t nz = op->colptr[t][jl]; nz < op->colptr[t][jl + 1]; nz++) {
      uint64_t row = (uint64_t)op->rowval[t][nz];
      cplx val = op->nzval[t][nz];
      /* change!(dgt, I, row) (AbstractBasis.jl:380-392) */
      uint64_t sn = s;
      uint64_t rr = row;
      for (int q = ns - 1; q >= 0; q--) {
        uint64_t w = b->pw[b->L - S[q]];
        uint64_t old = (sn / w) % b->base;
        uint64_t nd = rr % b->base;
        rr /= b->base;
        sn = sn - old * w + nd * w;
      }
      cplx cb;
      int64_t pos;
      basis_index(b, sn, &cb, &pos);
      for (int64_t q = 0; q < nb; q++) /* target[pos] += coeff * C * val */
        target[(pos - 1) * nb + q] += (C[q] * cb) * val;
    }
  }
}

/* Column j of the sector matrix as the reference builds it (colmn! with
 * coeff = 1, Operator.jl:458-465), returned as (1-based pos, value) pairs
 * merged by position in first-touch order.  Returns the number of pairs. */
int64_t orc_column(const orc_op *op, int64_t j, int64_t *pos, double *val, int64_t cap) {
  const orc_basis *b = op->b;
  int64_t n = 0;
  uint64_t s = content0(b, j);
  double r = basis_norm(b, j);
  cplx C = (r == 1.0) ? 1.0 : 1.0 / r;
  for (int64_t t = 0; t < op->nterms; t++) {
    int ns = op->nsite[t];
    const int *S = op->sites[t];
    uint64_t jl = 0;
    for (int q = 0; q < ns; q++) jl = jl * b->base + (s / b->pw[b->L - S[q]]) % b->base;
    for (int64_t nz = op->colptr[t][jl]; nz < op->colptr[t][jl + 1]; nz++) {
      uint64_t row = (uint64_t)op->rowval[t][nz], sn = s, rr = row;
      for (int q = ns - 1; q >= 0; q--) {
        uint64_t w = b->pw[b->L - S[q]];
        uint64_t old = (sn / w) % b->base;
        sn = sn - old * w + (rr % b->base) * w;
        rr /= b->base;
      }
      cplx cb;
      int64_t p;
      basis_index(b, sn, &cb, &p);
      if (cb == 0.0) continue;
      cplx v = (C * cb) * op->nzval[t][nz];
      int64_t q = 0;
      while (q < n && pos[q] != p) q++;
      if (q == n) {
        if (n == cap) return -1;
        pos[n] = p;
        val[2 * n] = 0;
        val[2 * n + 1] = 0;
        n++;
      }
      val[2 * q] += creal(v);
      val[2 * q + 1] += cimag(v);
    }
  }
  return n;
}

typedef struct {
  const orc_op *op;
  cplx *buf;
  const cplx *v;
  int64_t lo, hi, nb;
} mul_job;

static void *mul_worker(void *arg) {
  mul_job *m = (mul_job *)arg;
  for (int64_t j = m->lo; j < m->hi; j++) colmn(m->op, m->buf, j, m->v + j * m->nb, m->nb);
  return NULL;
}

void orc_mul_block(const orc_op *op, double *Y, const double *X, int64_t nb, int nthreads) {
  int64_t D = op->b->D;
  cplx *target = (cplx *)Y;
  const cplx *v = (const cplx *)X;
  if (nb < 1 || nb > 64) return;
  if (nthreads <= 1) { /* mul! (Operator.jl:309-315, 322-328) */
    for (int64_t j = 0; j < D; j++) colmn(op, target, j, v + j * nb, nb);
    return;
  }
  /* mul (Operator.jl:340-352): dividerange slices, private buffers, sum */
  int nt = nthreads;
  mul_job *jobs = (mul_job *)calloc(nt, sizeof(mul_job));
  pthread_t *th = (pthread_t *)calloc(nt, sizeof(pthread_t));
  int64_t each = D / nt, left = D % nt, start = 0;
  for (int t = 0; t < nt; t++) { /* dividerange (ProjectedBasis.jl:105-121) */
    int64_t cnt = each + (t < left ? 1 : 0);
    if (t == nt - 1) cnt = D - start;
    jobs[t].op = op;
    jobs[t].v = v;
    jobs[t].lo = start;
    jobs[t].hi = start + cnt;
    jobs[t].nb = nb;
    jobs[t].buf = (cplx *)calloc((size_t)(D ? D : 1) * nb, sizeof(cplx));
    start += cnt;
  }
  for (int t = 1; t < nt; t++) pthread_create(&th[t], NULL, mul_worker, &jobs[t]);
  mul_worker(&jobs[0]);
  for (int t = 1; t < nt; t++) pthread_join(th[t], NULL);
  for (int64_t i = 0; i < D * nb; i++) {
    cplx acc = jobs[0].buf[i];
    for (int t = 1; t < nt; t++) acc += jobs[t].buf[i];
    target[i] += acc;
  }
  for (int t = 0; t < nt; t++) free(jobs[t].buf);
  free(jobs);
  free(th);
}

typedef struct {
  const orc_op *op;
  cplx *buf;
  const cplx *v;
  const int64_t *cols;
  int64_t lo, hi;
} cols_job;

static void *cols_worker(void *arg) {
  cols_job *m = (cols_job *)arg;
  for (int64_t q = m->lo; q < m->hi; q++) colmn(m->op, m->buf, m->cols[q], m->v + m->cols[q], 1);
  return NULL;
}

/* The threaded mul (Operator.jl:340-352) restricted to a list of columns:
 * dividerange over the list, per-thread private full-length buffers, summed
 * onto target.  Used to time a bounded sample of a full matvec. */
void orc_mul_cols(const orc_op *op, double *target, const double *v, const int64_t *cols,
                  int64_t ncols, int nthreads) {
  int64_t D = op->b->D;
  int nt = nthreads < 1 ? 1 : nthreads;
  cols_job *jobs = (cols_job *)calloc(nt, sizeof(cols_job));
  pthread_t *th = (pthread_t *)calloc(nt, sizeof(pthread_t));
  int64_t each = ncols / nt, left = ncols % nt, start = 0;
  for (int t = 0; t < nt; t++) {
    int64_t cnt = each + (t < left ? 1 : 0);
    jobs[t].op = op;
    jobs[t].v = (const cplx *)v;
    jobs[t].cols = cols;
    jobs[t].lo = start;
    jobs[t].hi = start + cnt;
    jobs[t].buf = (cplx *)calloc((size_t)(D ? D : 1), sizeof(cplx));
    start += cnt;
  }
  for (int t = 1; t < nt; t++) pthread_create(&th[t], NULL, cols_worker, &jobs[t]);
  cols_worker(&jobs[0]);
  for (int t = 1; t < nt; t++) pthread_join(th[t], NULL);
  cplx *tg = (cplx *)target;
  for (int64_t i = 0; i < D; i++) {
    cplx acc = jobs[0].buf[i];
    for (int t = 1; t < nt; t++) acc += jobs[t].buf[i];
    tg[i] += acc;
  }
  for (int t = 0; t < nt; t++) free(jobs[t].buf);
  free(jobs);
  free(th);
}

void orc_mul(const orc_op *op, double *target, const double *v, int nthreads) {
  orc_mul_block(op, target, v, 1, nthreads);
}

typedef struct {
  const orc_op *op;
  int64_t lo, hi, count;
} cnt_job;

static void *cnt_worker(void *arg) {
  cnt_job *c = (cnt_job *)arg;
  const orc_op *op = c->op;
  const orc_basis *b = op->b;
  for (int64_t j = c->lo; j < c->hi; j++) {
    uint64_t s = content0(b, j);
    for (int64_t t = 0; t < op->nterms; t++) {
      int ns = op->nsite[t];
      const int *S = op->sites[t];
      uint64_t jl = 0;
      for (int q = 0; q < ns; q++) jl = jl * b->base + (s / b->pw[b->L - S[q]]) % b->base;
      for (int64_t nz = op->colptr[t][jl]; nz < op->colptr[t][jl + 1]; nz++) {
        uint64_t row = (uint64_t)op->rowval[t][nz];
        if (row == jl) continue;
        uint64_t sn = s, rr = row;
        for (int q = ns - 1; q >= 0; q--) {
          uint64_t w = b->pw[b->L - S[q]];
          uint64_t old = (sn / w) % b->base;
          sn = sn - old * w + (rr % b->base) * w;
          rr /= b->base;
        }
        cplx cb;
        int64_t pos;
        basis_index(b, sn, &cb, &pos);
        if (cb != 0.0) c->count++;
      }
    }
  }
  return NULL;
}

int64_t orc_count_offdiag(const orc_op *op, int nthreads) {
  int64_t D = op->b->D;
  int nt = nthreads < 1 ? 1 : nthreads;
  cnt_job *jobs = (cnt_job *)calloc(nt, sizeof(cnt_job));
  pthread_t *th = (pthread_t *)calloc(nt, sizeof(pthread_t));
  for (int t = 0; t < nt; t++) {
    jobs[t].op = op;
    jobs[t].lo = D * t / nt;
    jobs[t].hi = D * (t + 1) / nt;
  }
  for (int t = 1; t < nt; t++) pthread_create(&th[t], NULL, cnt_worker, &jobs[t]);
  cnt_worker(&jobs[0]);
  for (int t = 1; t < nt; t++) pthread_join(th[t], NULL);
  int64_t total = 0;
  for (int t = 0; t < nt; t++) total += jobs[t].count;
  free(jobs);
  free(th);
  return total;
}

/* ------------------------------------------------------------------ */
/* Symmetric tridiagonal eigensolver: implicit QL with Wilkinson shift  */
/* (the textbook method, Golub & Van Loan 8.3 / EISPACK tql2).          */
/* ------------------------------------------------------------------ */
int orc_tridiag_eig(int n, const double *alpha, const double *beta, double *d, double *Z) {
  double *e = (double *)calloc(n + 1, sizeof(double));
  for (int i = 0; i < n; i++) d[i] = alpha[i];
  for (int i = 0; i < n - 1; i++) e[i] = beta[i];
  for (int i = 0; i < n * n; i++) Z[i] = 0.0;
  for (int i = 0; i < n; i++) Z[i + i * n] = 1.0;
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < n; l++) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; m++) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= eps * dd) break;
      }
      if (m != l) {
        if (iter++ == 100) {
          free(e);
          set_err("tridiagonal eigensolver did not converge");
          return -1;
        }
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; i--) {
          double f = s * e[i], bb = c * e[i];
          r = hypot(f, g);
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
          for (int k = 0; k < n; k++) {
            double zf = Z[k + (i + 1) * n];
            Z[k + (i + 1) * n] = s * Z[k + i * n] + c * zf;
            Z[k + i * n] = c * Z[k + i * n] - s * zf;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  /* ascending order (selection sort keeps it simple for m <= a few 100) */
  for (int i = 0; i < n - 1; i++) {
    int k = i;
    for (int j = i + 1; j < n; j++)
      if (d[j] < d[k]) k = j;
    if (k != i) {
      double t = d[i];
      d[i] = d[k];
      d[k] = t;
      for (int q = 0; q < n; q++) {
        t = Z[q + i * n];
        Z[q + i * n] = Z[q + k * n];
        Z[q + k * n] = t;
      }
    }
  }
  free(e);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Krylov: TimeEvolution.jl                                             */
/* ------------------------------------------------------------------ */
static cplx zdot(const cplx *a, const cplx *b, int64_t n) {
  cplx s = 0.0;
  for (int64_t i = 0; i < n; i++) s += conj(a[i]) * b[i];
  return s;
}
static double znorm(const cplx *a, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) s += creal(a[i]) * creal(a[i]) + cimag(a[i]) * cimag(a[i]);
  return sqrt(s);
}

struct orc_krylov {
  const orc_op *H;
  int64_t N;
  int nthreads;
  double t_anchor, t_latest;
  cplx *psi_anchor;
  double norm_anchor;
  cplx *V; /* N x (m_max+1), column-major */
  double *alpha, *beta;
  int m;
  double *lambda, *Q, *Qe1, *Qem;
  double tol;
  int m_init, m_max, extend_step, nsample, bisect_iters, extend_basis, normalize_output;
  cplx *w;
  int64_t builds, extensions, restarts, matvecs, served, max_dim;
  int64_t n_accepted;
};

/* _lanczos_build! (TimeEvolution.jl:250-306) with _apply! (:237-241) */
static int lanczos_build(orc_krylov *c, int m_target) {
  int64_t N = c->N;
  if (m_target > c->m_max) m_target = c->m_max;
  for (int j = c->m + 1; j <= m_target; j++) {
    cplx *vj = c->V + (int64_t)(j - 1) * N;
    memset(c->w, 0, N * sizeof(cplx));
    orc_mul(c->H, (double *)c->w, (const double *)vj, c->nthreads);
    c->matvecs++;
    if (j > 1) {
      cplx *vjm1 = c->V + (int64_t)(j - 2) * N;
      for (int64_t i = 0; i < N; i++) c->w[i] -= c->beta[j - 2] * vjm1[i];
    }
    double aj = creal(zdot(vj, c->w, N));
    c->alpha[j - 1] = aj;
    for (int64_t i = 0; i < N; i++) c->w[i] -= aj * vj[i];
    for (int sweep = 0; sweep < 2; sweep++) {
      for (int i = 1; i <= j; i++) {
        cplx *vi = c->V + (int64_t)(i - 1) * N;
        cplx s = zdot(vi, c->w, N);
        for (int64_t k = 0; k < N; k++) c->w[k] -= s * vi[k];
      }
    }
    double bj = znorm(c->w, N);
    c->beta[j - 1] = bj;
    double sq = 1.4901161193847656e-08; /* sqrt(eps(Float64)) */
    if (bj <= sq * (fabs(aj) > 1.0 ? fabs(aj) : 1.0)) {
      c->m = j;
      c->beta[j - 1] = 0.0;
      return 1;
    }
    c->m = j;
    if (j + 1 <= c->m_max + 1) {
      cplx *vn = c->V + (int64_t)j * N;
      double inv = 1.0 / bj;
      for (int64_t i = 0; i < N; i++) vn[i] = c->w[i] * inv;
    }
  }
  return 0;
}

/* _build_reduced! (TimeEvolution.jl:311-323) */
static int build_reduced(orc_krylov *c) {
  int m = c->m;
  if (m < 1) {
    set_err("Cannot build reduced propagator for empty basis");
    return -1;
  }
  free(c->lambda);
  free(c->Q);
  free(c->Qe1);
  free(c->Qem);
  c->lambda = (double *)malloc(m * sizeof(double));
  c->Q = (double *)malloc((size_t)m * m * sizeof(double));
  c->Qe1 = (double *)malloc(m * sizeof(double));
  c->Qem = (double *)malloc(m * sizeof(double));
  if (orc_tridiag_eig(m, c->alpha, c->beta, c->lambda, c->Q)) return -1;
  for (int k = 0; k < m; k++) {
    c->Qe1[k] = c->Q[0 + k * m];
    c->Qem[k] = c->Q[(m - 1) + k * m];
  }
  return 0;
}

static inline cplx cis(double x) { return cos(x) + IU * sin(x); }

/* _defect_monitor (TimeEvolution.jl:341-350) */
static double defect(const orc_krylov *c, double tau) {
  int m = c->m;
  double bm = c->beta[m - 1];
  if (bm == 0.0) return 0.0;
  cplx acc = 0.0;
  for (int k = 0; k < m; k++) acc += (c->Qem[k] * cis(-tau * c->lambda[k])) * c->Qe1[k];
  return fabs(bm) * cabs(acc);
}

/* _reconstruct_state! (TimeEvolution.jl:352-362) with _reduced_coeffs (:329-338) */
static void reconstruct(const orc_krylov *c, double tau, cplx *out) {
  int m = c->m;
  cplx *ph = (cplx *)malloc(m * sizeof(cplx)), *cc = (cplx *)malloc(m * sizeof(cplx));
  for (int k = 0; k < m; k++) ph[k] = cis(-tau * c->lambda[k]) * c->Qe1[k];
  for (int i = 0; i < m; i++) {
    cplx s = 0.0;
    for (int k = 0; k < m; k++) s += c->Q[i + k * m] * ph[k];
    cc[i] = s;
  }
  for (int64_t r = 0; r < c->N; r++) {
    cplx s = 0.0;
    for (int k = 0; k < m; k++) s += c->V[r + (int64_t)k * c->N] * cc[k];
    out[r] = s * c->norm_anchor;
  }
  if (c->normalize_output) {
    double nrm = znorm(out, c->N);
    if (nrm > 0)
      for (int64_t r = 0; r < c->N; r++) out[r] /= nrm;
  }
  free(ph);
  free(cc);
}

/* _effective_nsample (TimeEvolution.jl:383-394) */
static int eff_nsample(const orc_krylov *c, double tau) {
  int base = c->nsample;
  if (c->m <= 1) return base;
  double om = c->lambda[c->m - 1] - c->lambda[0];
  if (om <= 0) return base;
  int nn = (int)ceil(4.0 * om * tau / M_PI) + 1;
  int v = nn > base ? nn : base;
  if (v > 513) v = 513;
  return v;
}

/* _max_valid_interval (TimeEvolution.jl:396-429) */
static double max_valid(const orc_krylov *c, double tau) {
  if (!(tau > 0)) return 0.0;
  if (fabs(c->beta[c->m - 1]) <= c->tol) return tau;
  int n = eff_nsample(c, tau);
  double prev = 0.0;
  for (int k = 1; k <= n; k++) {
    double x = (n == 1) ? 0.0 : tau * (double)(k - 1) / (double)(n - 1);
    double eta = defect(c, x);
    if (!isfinite(eta) || eta > c->tol) {
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

/* _extend_basis! (TimeEvolution.jl:434-447) */
static int extend(orc_krylov *c) {
  if (!c->extend_basis) return 0;
  if (c->m >= c->m_max) return 0;
  int old = c->m;
  int nm = c->m + c->extend_step < c->m_max ? c->m + c->extend_step : c->m_max;
  lanczos_build(c, nm);
  if (c->m > old) {
    build_reduced(c);
    c->extensions++;
    if (c->m > c->max_dim) c->max_dim = c->m;
    return 1;
  }
  return 0;
}

/* _restart_from_new_anchor! (TimeEvolution.jl:449-477) */
static int restart(orc_krylov *c, double tau) {
  cplx *pn = (cplx *)malloc(c->N * sizeof(cplx));
  reconstruct(c, tau, pn);
  c->t_anchor += tau;
  memcpy(c->psi_anchor, pn, c->N * sizeof(cplx));
  free(pn);
  double nrm = znorm(c->psi_anchor, c->N);
  if (!(nrm > 0)) {
    set_err("Restart failed: anchor state has zero norm");
    return -1;
  }
  c->norm_anchor = nrm;
  memset(c->alpha, 0, c->m_max * sizeof(double));
  memset(c->beta, 0, c->m_max * sizeof(double));
  c->m = 0;
  for (int64_t i = 0; i < c->N; i++) c->V[i] = c->psi_anchor[i] / c->norm_anchor;
  lanczos_build(c, c->m_init);
  build_reduced(c);
  c->builds++;
  c->restarts++;
  if (c->m > c->max_dim) c->max_dim = c->m;
  return 0;
}

/* _ensure_valid_up_to! (TimeEvolution.jl:486-520) */
static int ensure_valid(orc_krylov *c, double target) {
  double rem = target;
  if (rem <= 0) return 0;
  while (1) {
    double v = max_valid(c, rem);
    if (v >= rem) return 0;
    if (c->extend_basis && c->m < c->m_max)
      if (extend(c)) continue;
    if (v <= 0) {
      if (c->m < c->m_max && extend(c)) continue;
      set_err("KrylovEvolutionCache: basis is not valid at tau=0; consider increasing m_max or loosening tol");
      return -1;
    }
    c->n_accepted++;
    if (restart(c, v)) return -1;
    rem -= v;
  }
}

/* KrylovEvolutionCache(H, psi0; ...) (TimeEvolution.jl:166-218) */
orc_krylov *orc_krylov_create(const orc_op *op, const double *psi0, double tol,
                              int m_init, int m_max, int extend_step,
                              int nsample, int bisect_iters, int extend_basis,
                              int normalize_output, int nthreads) {
  int64_t N = op->b->D;
  if (N <= 0) { set_err("Initial state must be non-empty"); return NULL; }
  if (m_init < 1) { set_err("m_init must be >= 1"); return NULL; }
  if (m_max < m_init) { set_err("m_max must be >= m_init"); return NULL; }
  if (m_max > N) m_max = (int)N;
  if (m_init > m_max) m_init = m_max;
  if (nsample < 2) { set_err("nsample must be >= 2"); return NULL; }
  if (bisect_iters < 0) { set_err("bisect_iters must be >= 0"); return NULL; }
  if (extend_step < 1) { set_err("extend_step must be >= 1"); return NULL; }
  orc_krylov *c = (orc_krylov *)calloc(1, sizeof(orc_krylov));
  c->H = op;
  c->N = N;
  c->nthreads = nthreads;
  c->tol = tol;
  c->m_init = m_init;
  c->m_max = m_max;
  c->extend_step = extend_step;
  c->nsample = nsample;
  c->bisect_iters = bisect_iters;
  c->extend_basis = extend_basis;
  c->normalize_output = normalize_output;
  c->psi_anchor = (cplx *)malloc(N * sizeof(cplx));
  memcpy(c->psi_anchor, psi0, N * sizeof(cplx));
  double nrm = znorm(c->psi_anchor, N);
  if (!(nrm > 0)) {
    set_err("Initial state has zero norm");
    free(c->psi_anchor);
    free(c);
    return NULL;
  }
  c->norm_anchor = nrm;
  c->V = (cplx *)calloc((size_t)N * (m_max + 1), sizeof(cplx));
  for (int64_t i = 0; i < N; i++) c->V[i] = c->psi_anchor[i] / nrm;
  c->alpha = (double *)calloc(m_max, sizeof(double));
  c->beta = (double *)calloc(m_max, sizeof(double));
  c->w = (cplx *)calloc(N, sizeof(cplx));
  lanczos_build(c, c->m_init);
  build_reduced(c);
  c->builds = 1;
  c->max_dim = c->m;
  return c;
}

/* timeevolve!(cache, t) (TimeEvolution.jl:586-597) */
int orc_krylov_evolve(orc_krylov *c, double t, double *out) {
  if (t < c->t_latest) {
    set_err("timeevolve!: requested time precedes the last served time; backward evolution is not supported");
    return -1;
  }
  double tau = t - c->t_anchor;
  if (tau < 0) {
    set_err("timeevolve!: requested time precedes anchor time");
    return -1;
  }
  if (ensure_valid(c, tau)) return -1;
  reconstruct(c, t - c->t_anchor, (cplx *)out);
  c->t_latest = t;
  c->served++;
  return 0;
}

void orc_krylov_diagnostics(const orc_krylov *c, int64_t *o) {
  o[0] = c->builds;
  o[1] = c->extensions;
  o[2] = c->restarts;
  o[3] = c->matvecs;
  o[4] = c->served;
  o[5] = c->max_dim;
  o[6] = c->n_accepted;
}
int orc_krylov_m(const orc_krylov *c) { return c->m; }

void orc_krylov_free(orc_krylov *c) {
  if (!c) return;
  free(c->psi_anchor);
  free(c->V);
  free(c->alpha);
  free(c->beta);
  free(c->w);
  free(c->lambda);
  free(c->Q);
  free(c->Qe1);
  free(c->Qem);
  free(c);
}

/* Lanczos ground state: the _lanczos_build! step (TimeEvolution.jl:250-306)
 * followed after every step by the lowest Ritz pair of T_j and its residual
 * |beta_j Q[j,1]| (the same algebra as _defect_monitor :341-350). */
int orc_lanczos_ground(const orc_op *op, const double *v0, int m, double tol,
                       int nthreads, double *E0, double *resid, double *vec) {
  int64_t N = op->b->D;
  if (m > N) m = (int)N;
  orc_krylov c;
  memset(&c, 0, sizeof(c));
  c.H = op;
  c.N = N;
  c.nthreads = nthreads;
  c.m_max = m;
  c.V = (cplx *)calloc((size_t)N * (m + 1), sizeof(cplx));
  c.alpha = (double *)calloc(m, sizeof(double));
  c.beta = (double *)calloc(m, sizeof(double));
  c.w = (cplx *)calloc(N, sizeof(cplx));
  const cplx *x0 = (const cplx *)v0;
  double nrm = znorm(x0, N);
  if (!(nrm > 0)) {
    set_err("Initial state has zero norm");
    free(c.V); free(c.alpha); free(c.beta); free(c.w);
    return -1;
  }
  for (int64_t i = 0; i < N; i++) c.V[i] = x0[i] / nrm;
  double *lam = (double *)malloc(m * sizeof(double));
  double *Q = (double *)malloc((size_t)m * m * sizeof(double));
  int steps = 0;
  double e0 = 0, res = 0;
  for (int j = 1; j <= m; j++) {
    int broke = lanczos_build(&c, j);
    steps = c.m;
    orc_tridiag_eig(steps, c.alpha, c.beta, lam, Q);
    e0 = lam[0];
    res = broke ? 0.0 : fabs(c.beta[steps - 1] * Q[(steps - 1) + 0 * steps]);
    if (broke || (tol > 0 && res <= tol)) break;
  }
  *E0 = e0;
  *resid = res;
  if (vec) {
    cplx *o = (cplx *)vec;
    for (int64_t r = 0; r < N; r++) {
      cplx s = 0.0;
      for (int k = 0; k < steps; k++) s += c.V[r + (int64_t)k * N] * Q[k + 0 * steps];
      o[r] = s;
    }
  }
  free(lam); free(Q); free(c.V); free(c.alpha); free(c.beta); free(c.w);
  return steps;
}
