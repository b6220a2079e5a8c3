// edk_basis.hip -- sector enumeration on the device and the basis half of the
// C ABI (build / upload / download / index).
//
// Enumeration: the candidate space (fixed-popcount states unranked from the
// combinatorial number system for base 2 with N, else every content in
// [0, base^L) with the digit-sum filter) is cut into contiguous per-thread
// chunks.  Pass 1 runs the reference's orbit judge on every candidate and
// counts; an exclusive scan over thread and block counts gives offsets; pass 2
// re-runs the judge and writes.  Chunks are in increasing candidate order, so
// the table comes out sorted, like the reference's vcat of per-thread ranges
// (TranslationalBasis.jl:194-204).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "edk_internal.h"

// ---------------------------------------------------------------------------
// judges (exact reference control flow)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 edk_abelian_phase(const CanonDev &C, const int *os) {
  double2 acc = C.ph[C.gdesc[6] + os[0] - 1];
  for (int g = 1; g < C.ngen; g++) acc = cmul(acc, C.ph[C.gdesc[EDK_GDESC * g + 6] + os[g] - 1]);
  return acc;
}

// Returns 1 when s is a representative; *nrm gets its norm.
__device__ int edk_judge(const CanonDev &C, uint64_t s, double *nrm) {
  int tot = C.L * C.w, sh = C.A * C.w;
  if (C.kind == EDK_PROJECTED) {
    *nrm = 1.0;
    return 1;
  }
  if (C.kind == EDK_TRANSLATIONAL) {  // TranslationJudge, TranslationalBasis.jl:94-114
    uint64_t st = s;
    for (int n = 1; n < C.len; n++) {
      st = edk_rotr(st, sh, tot, C.mask);
      if (st < s) return 0;
      if (st == s) {
        if (((int64_t)n * C.kp) % C.len) return 0;
        *nrm = C.Rtab[n - 1];
        return 1;
      }
    }
    *nrm = C.Rtab[C.len - 1];
    return 1;
  }
  if (C.kind == EDK_TRANSLATION_PARITY) {  // TranslationalParityBasis.jl:68-112
    int R = C.len;
    bool qp = false;
    uint64_t In = edk_revfields(s, C.L, C.w);
    if (In < s) return 0;
    if (In == s) {
      qp = true;
      if (C.P != 1) return 0;
    }
    uint64_t st = s;
    for (int n = 1; n < C.len; n++) {
      st = edk_rotr(st, sh, tot, C.mask);
      uint64_t In2 = edk_revfields(st, C.L, C.w);
      if (st < s) return 0;
      if (In2 < s) return 0;
      bool evn = (n & 1) == 0;
      if (In2 == s) {
        int Kn = (C.kp == 1 || evn) ? 1 : -1;
        if (C.P * Kn != 1) return 0;
        qp = true;
      }
      if (st == s) {
        if (!(C.kp == 1 || evn)) return 0;
        R = n;
        break;
      }
    }
    *nrm = C.Rtab[(qp ? C.len : 0) + R - 1];
    return 1;
  }
  if (C.kind == EDK_PARITY || C.kind == EDK_FLIP) {  // ParityBasis.jl:44-57, FlipBasis.jl:42-55
    const uint64_t In = C.kind == EDK_PARITY ? edk_revfields(s, C.L, C.w) : edk_flipfields(C, s);
    if (In < s) return 0;
    if (In == s) {
      if (C.P != 1) return 0;
      *nrm = C.Rtab[1];
    } else {
      *nrm = C.Rtab[0];
    }
    return 1;
  }
  if (C.kind == EDK_PARITY_FLIP) {  // double_parity_check, ParityFlipBasis.jl:47-65, 73-79
    const uint64_t Ip = edk_revfields(s, C.L, C.w);
    if (Ip < s) return 0;
    const uint64_t Iz = edk_flipfields(C, s);
    if (Iz < s) return 0;
    const uint64_t Ipz = edk_flipfields(C, Ip);
    if (Ipz < s) return 0;
    int N = 1;
    if (Ip == s) {
      if (C.P != 1) return 0;
      N++;
    }
    if (Iz == s) {
      if (C.Z != 1) return 0;
      N++;
    }
    if (Ipz == s) {
      if (C.P * C.Z != 1) return 0;
      N++;
    }
    *nrm = C.Rtab[N - 1];
    return 1;
  }
  if (C.kind == EDK_TRANSLATION_FLIP) {  // TranslationFlipJudge, TranslationalFlipBasis.jl:50-94
    int R = C.len;
    bool qp = false;
    const uint64_t In = edk_flipfields(C, s);
    if (In < s) return 0;
    if (In == s) {
      qp = true;
      if (C.P != 1) return 0;
    }
    const int half = C.len / 2;
    uint64_t st = s;
    for (int n = 1; n < C.len; n++) {
      st = edk_rotr(st, sh, tot, C.mask);  // circshift!(dgt, A)
      const uint64_t In2 = edk_flipfields(C, st);
      if (st < s) return 0;
      if (In2 < s) return 0;
      if (In2 == s) {
        // isequal(isone(P), iseven(div(n * K, L / 2)))
        if ((C.P == 1) != (((n * C.kp) / half) % 2 == 0)) return 0;
        qp = true;
      }
      if (st == s) {
        if ((n * C.kp) % C.len) return 0;
        R = n;
        break;
      }
    }
    *nrm = C.Rtab[(qp ? C.len : 0) + R - 1];
    return 1;
  }
  // check_min_int / check_min, AbelianBasis.jl:322-338, 399-418
  int os[EDK_MAXGEN];
  for (int g = 0; g < C.ngen; g++) os[g] = 1;
  uint64_t cur = s;
  int N = 1;
  for (int step = 2; step <= C.order; step++) {
    for (int g = 0; g < C.ngen; g++) {
      cur = edk_apply_gen(C, g, cur);
      if (++os[g] > C.gdesc[EDK_GDESC * g + 1]) os[g] = 1;
      else break;
    }
    if (cur < s) return 0;
    if (cur == s) {
      double2 ph = edk_abelian_phase(C, os);
      if (!(hypot(ph.x - 1.0, ph.y) < 1e-14)) return 0;
      N++;
    }
  }
  *nrm = C.Rtab[N - 1];
  return 1;
}

// ---------------------------------------------------------------------------
// candidate generation
// ---------------------------------------------------------------------------
struct EnumArgs {
  CanonDev C;
  int mode;  // 0: fixed-popcount ranks (base 2), 1: content scan
  int L, base, w, pop, dsum;
  uint64_t total, per_thread;
  const uint64_t *binom;  // (EDK_MAXL+1)^2, binom[n*(EDK_MAXL+1)+k]
  const uint64_t *pw;     // base^i
};

__device__ __forceinline__ uint64_t edk_unrank(uint64_t r, int L, int n, const uint64_t *binom) {
  uint64_t s = 0;
  for (int pos = L - 1; pos >= 0 && n > 0; pos--) {
    uint64_t c = binom[pos * (EDK_MAXL + 1) + n];
    if (r >= c) {
      s |= 1ull << pos;
      r -= c;
      n--;
    }
  }
  return s;
}

// _gosper_next (AbelianBasis.jl:456-463)
__device__ __forceinline__ uint64_t edk_gosper_next(uint64_t s) {
  uint64_t c = s & (~s + 1);
  uint64_t r = s + c;
  return r | (((r ^ s) >> 2) / c);
}

__device__ __forceinline__ uint64_t edk_content_to_packed(uint64_t c, int L, int base, int w) {
  if (w == 1 && base == 2) return c;
  uint64_t s = 0;
  for (int q = 0; q < L; q++) {
    s |= (c % (uint64_t)base) << (q * w);
    c /= (uint64_t)base;
  }
  return s;
}

__device__ __forceinline__ int edk_digit_sum(uint64_t s, int L, int w) {
  if (w == 1) return __popcll(s);
  if (w == 2)
    return __popcll(s & 0x5555555555555555ull) + 2 * __popcll(s & 0xAAAAAAAAAAAAAAAAull);
  int t = 0;
  uint64_t fm = (1ull << w) - 1;
  for (int q = 0; q < L; q++) t += (int)((s >> (q * w)) & fm);
  return t;
}

// packed + 1 in base-B digit arithmetic
__device__ __forceinline__ uint64_t edk_packed_inc(uint64_t s, int L, int base, int w) {
  uint64_t fm = (1ull << w) - 1;
  for (int q = 0; q < L; q++) {
    uint64_t d = (s >> (q * w)) & fm;
    if (d + 1 < (uint64_t)base) return s + (1ull << (q * w));
    s &= ~(fm << (q * w));
  }
  return s;
}

#define ENUM_TPB 256

__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t *total) {
  __shared__ uint32_t wsum[ENUM_TPB / 64];
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  uint32_t base = 0, tot = 0;
  for (int q = 0; q < ENUM_TPB / 64; q++) {
    if (q < wid) base += wsum[q];
    tot += wsum[q];
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

template <bool WRITE>
__global__ void __launch_bounds__(ENUM_TPB) k_enum(EnumArgs A, uint64_t *blockcount,
                                                   const uint64_t *blockoff, uint64_t *keys,
                                                   double *R) {
  uint64_t t = (uint64_t)blockIdx.x * ENUM_TPB + threadIdx.x;
  uint64_t lo = t * A.per_thread;
  uint64_t hi = lo + A.per_thread;
  if (hi > A.total) hi = A.total;
  uint32_t cnt = 0;
  double nrm;
  uint64_t s = 0;
  if (lo < hi) s = A.mode == 0 ? edk_unrank(lo, A.L, A.pop, A.binom)
                               : edk_content_to_packed(lo, A.L, A.base, A.w);
  uint64_t s0 = s;
  for (uint64_t q = lo; q < hi; q++) {
    bool cand = A.mode == 0 || A.dsum < 0 || edk_digit_sum(s, A.L, A.w) == A.dsum;
    if (cand && edk_judge(A.C, s, &nrm)) cnt++;
    if (q + 1 < hi) s = A.mode == 0 ? edk_gosper_next(s) : edk_packed_inc(s, A.L, A.base, A.w);
  }
  uint32_t total;
  uint32_t off = block_excl_scan_u32(cnt, &total);
  if (!WRITE) {
    if (threadIdx.x == 0) blockcount[blockIdx.x] = total;
    return;
  }
  uint64_t o = blockoff[blockIdx.x] + off;
  s = s0;
  for (uint64_t q = lo; q < hi; q++) {
    bool cand = A.mode == 0 || A.dsum < 0 || edk_digit_sum(s, A.L, A.w) == A.dsum;
    if (cand && edk_judge(A.C, s, &nrm)) {
      keys[o] = s;
      R[o] = nrm;
      o++;
    }
    if (q + 1 < hi) s = A.mode == 0 ? edk_gosper_next(s) : edk_packed_inc(s, A.L, A.base, A.w);
  }
}

// exclusive scan of n uint64 counts into off[0..n] (off[n] = total); one block
__global__ void __launch_bounds__(1024) k_scan_counts(const uint64_t *cnt, int64_t n, uint64_t *off) {
  __shared__ uint64_t part[1024];
  int64_t per = (n + 1023) / 1024;
  int64_t lo = (int64_t)threadIdx.x * per, hi = lo + per;
  if (hi > n) hi = n;
  uint64_t s = 0;
  for (int64_t i = lo; i < hi; i++) s += cnt[i];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t acc = 0;
    for (int i = 0; i < 1024; i++) {
      uint64_t v = part[i];
      part[i] = acc;
      acc += v;
    }
    off[n] = acc;
  }
  __syncthreads();
  uint64_t acc = part[threadIdx.x];
  for (int64_t i = lo; i < hi; i++) {
    off[i] = acc;
    acc += cnt[i];
  }
}

// largest directory slot (integer maximum: order-independent)
__global__ void k_dir_maxgap(const uint32_t *dir, int64_t ndir, unsigned int *out) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned int g = b < ndir ? dir[b + 1] - dir[b] : 0u;
  for (int o = 32; o > 0; o >>= 1) g = max(g, (unsigned int)__shfl_down(g, o, 64));
  if ((threadIdx.x & 63) == 0 && g) atomicMax(out, g);
}

// directory: dir[b] = lower_bound(keys, b << dshift), dir[ndir] = D
__global__ void k_build_dir(const uint64_t *keys, int64_t D, int dshift, int64_t ndir, uint32_t *dir) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > ndir) return;
  if (b == ndir) {
    dir[b] = (uint32_t)D;
    return;
  }
  uint64_t key = (uint64_t)b << dshift;
  int64_t lo = 0, hi = D;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  dir[b] = (uint32_t)lo;
}

__global__ void k_index(CanonDev C, SearchDev S, const double *R, int kind, int L, int base, int w,
                        const int64_t *contents, int64_t n, int64_t *pos, double *cre, double *cim) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  uint64_t c = (uint64_t)contents[q];
  if (kind == EDK_TENSOR) {
    pos[q] = (int64_t)c + 1;
    cre[q] = 1.0;
    cim[q] = 0.0;
    return;
  }
  uint64_t s = edk_content_to_packed(c, L, base, w);
  CanonOut o = edk_canon(C, s);
  int64_t j = edk_search(S, o.rep);
  if (j < 0) {
    pos[q] = 1;
    cre[q] = 0.0;
    cim[q] = 0.0;
    return;
  }
  pos[q] = j + 1;
  double r = R[j];
  cre[q] = o.phase.x * r;
  cim[q] = o.phase.y * r;
}

__global__ void k_packed_to_content(const uint64_t *keys, int64_t D, int L, int base, int w, int64_t *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D) return;
  uint64_t s = keys[i];
  uint64_t fm = (1ull << w) - 1, c = 0;
  for (int q = L - 1; q >= 0; q--) c = c * (uint64_t)base + ((s >> (q * w)) & fm);
  out[i] = (int64_t)c + 1;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
CanonDev edk_basis::canon() const {
  CanonDev c;
  c.kind = kind;
  c.L = L;
  c.w = w;
  c.A = A;
  c.len = len;
  c.base = base;
  int tot = L * w;
  c.mask = tot >= 64 ? ~0ull : ((1ull << tot) - 1);
  uint64_t lf = 0;
  for (int q = 0; q < L; q++) lf |= 1ull << (q * w);
  c.lowfields = lf;
  c.kp = kp;
  c.P = P;
  c.Z = Z;
  c.ngen = ngen;
  c.order = order;
  c.Rtab = d_Rtab;
  c.ph = d_ph;
  c.gdesc = d_gdesc;
  c.gdata = d_gdata;
  c.gshift = d_gshift;
  c.ptab = d_ptab;
  c.pbits = pbits;
  return c;
}

SearchDev edk_basis::search() const {
  SearchDev s;
  s.keys = d_keys;
  s.D = D;
  s.dir = d_dir;
  s.dshift = dshift;
  s.ndir = ndir;
  s.nprobe = nprobe;
  return s;
}

static uint64_t g_binom[(EDK_MAXL + 1) * (EDK_MAXL + 1)];
static void binom_init() {
  if (g_binom[0]) return;
  for (int n = 0; n <= EDK_MAXL; n++) {
    g_binom[n * (EDK_MAXL + 1)] = 1;
    for (int k = 1; k <= n; k++)
      g_binom[n * (EDK_MAXL + 1) + k] =
          g_binom[(n - 1) * (EDK_MAXL + 1) + k - 1] + (k <= n - 1 ? g_binom[(n - 1) * (EDK_MAXL + 1) + k] : 0);
  }
}

// compile_benes (AbelianBasis.jl:28-69) on w-bit fields
static void benes_fields(const int32_t *perm, int L, int w, std::vector<uint64_t> &mask,
                         std::vector<int> &shift) {
  std::vector<int> bit_perm(L), cur(L), where(L), inv(L);
  for (int i = 1; i <= L; i++) bit_perm[L - i] = L - perm[i - 1];
  for (int i = 0; i < L; i++) cur[i] = where[i] = i;
  for (int s = 0; s < L; s++) inv[bit_perm[s]] = s;
  uint64_t fm = (1ull << w) - 1;
  for (int dst = 0; dst < L; dst++) {
    int src = inv[dst], cp = where[src];
    if (cp != dst) {
      int lo = std::min(dst, cp), hi = std::max(dst, cp);
      mask.push_back(fm << (lo * w));
      shift.push_back((hi - lo) * w);
      int other = cur[dst];
      std::swap(cur[dst], cur[cp]);
      where[src] = dst;
      where[other] = cp;
    }
  }
}

// _compute_period (AbelianBasis.jl:109-131)
static int compute_period(const int32_t *perm, const uint8_t *inv, int L) {
  std::vector<int> pos(L), np(L);
  std::vector<uint8_t> fl(L, 0), nf(L);
  for (int i = 0; i < L; i++) pos[i] = i + 1;
  int lim = std::max(2 * L * L, 100);
  for (int p = 1; p <= lim; p++) {
    bool ident = true;
    for (int i = 0; i < L; i++) {
      np[i] = perm[pos[i] - 1];
      nf[i] = fl[i] ^ (inv ? inv[pos[i] - 1] : 0);
    }
    for (int i = 0; i < L; i++) {
      pos[i] = np[i];
      fl[i] = nf[i];
      if (pos[i] != i + 1 || fl[i]) ident = false;
    }
    if (ident) return p;
  }
  return -1;
}

static void free_basis_mem(edk_basis *b) {
  hipFree(b->d_keys);
  hipFree(b->d_R);
  hipFree(b->d_dir);
  hipFree(b->d_Rtab);
  hipFree(b->d_ph);
  hipFree(b->d_gdesc);
  hipFree(b->d_gdata);
  hipFree(b->d_gshift);
  hipFree(b->d_ptab);
}

template <typename T>
static int upload_vec(edk_ctx *ctx, const std::vector<T> &h, T **d) {
  size_t n = std::max<size_t>(h.size(), 1);
  EDK_HIP(hipMalloc((void **)d, n * sizeof(T)));
  if (!h.empty())
    EDK_HIP(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return EDK_OK;
}

// Add one Abelian generator: AbelianOperator(order, k, perm; inv) (AbelianBasis.jl:167-193)
static int add_generator(edk_basis *b, int order, int k, const int32_t *perm, const uint8_t *inv) {
  if (b->ngen >= EDK_MAXGEN) return edk_fail(EDK_ERR_UNSUPPORTED, "too many symmetry generators");
  int L = b->L, w = b->w;
  int type = GEN_BENES, param = 0;
  bool ident = true, rev = true;
  for (int i = 1; i <= L; i++) {
    if (perm[i - 1] != i) ident = false;
    if (perm[i - 1] != L - i + 1) rev = false;
  }
  int rot = -1;
  for (int a = 1; a < L && rot < 0; a++) {
    bool ok = true;
    for (int i = 1; i <= L && ok; i++) ok = perm[i - 1] == ((i - a - 1) % L + L) % L + 1;
    if (ok) rot = a;
  }
  std::vector<uint64_t> lm;
  std::vector<int> ls;
  // the fields' bit displacements (site i sits at bit (L - i) w; its content
  // moves to site perm[i]): few distinct ones (a torus translation along the
  // fastest axis has two) give a shift-and-mask form
  std::map<int, uint64_t> disp;
  const uint64_t fmask = (1ull << w) - 1;
  for (int i = 1; i <= L; i++) {
    const int src = (L - i) * w, dst = (L - perm[i - 1]) * w;
    disp[dst - src] |= fmask << src;
  }
  if (ident) type = GEN_IDENT;
  else if (rev) type = GEN_REV;
  else if (rot > 0) {
    type = GEN_ROTL;  // content of site i moves to site i-a: a left rotation
    param = rot * w;
  } else if ((int)disp.size() <= EDK_MAXSHIFTS) {
    type = GEN_SHIFTS;
    for (auto &kv : disp) {
      lm.push_back(kv.second);
      ls.push_back(kv.first);
    }
  } else {
    benes_fields(perm, L, w, lm, ls);
  }
  uint64_t invmask = 0;
  if (inv) {
    uint64_t fm = (b->base == 3) ? 2ull : ((1ull << w) - 1);
    for (int i = 1; i <= L; i++)
      if (inv[i - 1]) invmask |= fm << ((L - i) * w);
    if (invmask && b->base > 4) return edk_fail(EDK_ERR_UNSUPPORTED, "spin inversion needs base <= 4 on the device");
  }
  int layer_off = (int)b->h_gdata.size();
  for (size_t q = 0; q < lm.size(); q++) {
    b->h_gdata.push_back(lm[q]);
    b->h_gshift.push_back(ls[q]);
  }
  int inv_slot = (int)b->h_gdata.size();
  b->h_gdata.push_back(invmask);
  b->h_gshift.push_back(0);
  int ph_off = (int)b->h_ph.size();
  for (int j = 0; j < order; j++) {
    double2 c;
    if (k == 0) c = make_double2(1.0, 0.0);
    else if (2 * k == order) c = make_double2((j % 2 == 0) ? 1.0 : -1.0, 0.0);
    else {
      double th = ((2.0 * M_PI * k) / order) * j;
      c = make_double2(cos(th), sin(th));
    }
    b->h_ph.push_back(c);
  }
  if (b->ngen == 0) {
    b->h_perm0.assign(perm, perm + L);
    b->h_inv0.assign(L, 0);
    if (inv) b->h_inv0.assign(inv, inv + L);
  }
  int d[EDK_GDESC] = {type, order, param, inv_slot, layer_off, (int)lm.size(), ph_off, 0};
  for (int q = 0; q < EDK_GDESC; q++) b->h_gdesc.push_back(d[q]);
  b->ngen++;
  return EDK_OK;
}

// Table of the pruned Abelian walk (edk_canon_abelian_pruned): applies with
// two or three generators, none a Benes network, generator 0 of order <= 32
// mapping the sites 1..k (the top k fields) onto themselves, for the largest
// such k with k w <= 10 (at most 1024 entries).  Generator 0 is applied here
// as the device applies it: fields move to their permuted sites, then the
// inversion mask.
static void build_prune_table(edk_basis *b, std::vector<uint2> &tab) {
  tab.clear();
  b->pbits = 0;
  if (b->kind != EDK_ABELIAN || b->ngen < 2 || b->ngen > 3 || (int)b->h_perm0.size() != b->L) return;
  for (int g = 0; g < b->ngen; g++)
    if (b->h_gdesc[EDK_GDESC * g] == GEN_BENES) return;
  const int L = b->L, w = b->w, n0 = b->h_gdesc[1];
  if (n0 < 1 || n0 > 32) return;
  int k = 0;
  for (int kk = 1; kk < L && kk * w <= 10; kk++) {
    bool inv_ok = true;
    for (int i = 1; i <= kk && inv_ok; i++) inv_ok = b->h_perm0[i - 1] >= 1 && b->h_perm0[i - 1] <= kk;
    if (inv_ok) k = kk;
  }
  if (k == 0) return;
  const int tot = L * w, bits = k * w, sh = tot - bits;
  const uint64_t fm = (1ull << w) - 1;
  uint64_t invmask = 0, lowf = 0;
  for (int i = 1; i <= L; i++) {
    lowf |= 1ull << ((L - i) * w);
    if (b->h_inv0[i - 1]) invmask |= ((b->base == 3) ? 2ull : fm) << ((L - i) * w);
  }
  auto apply0 = [&](uint64_t s) {
    uint64_t r = 0;
    for (int i = 1; i <= L; i++) r |= ((s >> ((L - i) * w)) & fm) << ((L - b->h_perm0[i - 1]) * w);
    if (invmask) {
      if (b->base != 3) r ^= invmask;
      else r ^= invmask & ~((r & lowf) << 1);
    }
    return r;
  };
  tab.assign((size_t)1 << bits, make_uint2(0, 0));
  for (uint64_t v = 0; v < ((uint64_t)1 << bits); v++) {
    uint64_t c = v << sh;
    uint32_t mn = 0xffffffffu, mask = 0;
    for (int a0 = 0; a0 < n0; a0++) {
      const uint32_t top = (uint32_t)(c >> sh);
      if (top < mn) {
        mn = top;
        mask = 0;
      }
      if (top == mn) mask |= 1u << a0;
      c = apply0(c);
    }
    tab[v] = make_uint2(mn, mask);
  }
  b->pbits = bits;
}

// Validate the spec and fill the canonicaliser tables (host + device).
static int setup_basis(edk_ctx *ctx, const edk_basis_spec *spec, edk_basis *b) {
  b->ctx = ctx;
  b->spec = *spec;
  b->spec.perms = nullptr;
  b->spec.ks = nullptr;
  b->spec.invs = nullptr;
  b->kind = spec->kind;
  b->L = spec->L;
  b->base = spec->base;
  b->A = spec->a > 0 ? spec->a : 1;
  if (b->L < 1 || b->L > EDK_MAXL || b->base < 2) return edk_fail(EDK_ERR_ARG, "invalid L or base");
  b->w = 1;
  while ((1 << b->w) < b->base) b->w++;
  if (b->L * b->w > 64) return edk_fail(EDK_ERR_UNSUPPORTED, "state does not fit a 64-bit packed word");
  b->pw.assign(b->L + 1, 1);
  long double lim = 1.0L;
  for (int i = 1; i <= b->L; i++) {
    b->pw[i] = b->pw[i - 1] * (uint64_t)b->base;
    lim *= b->base;
  }
  if (lim > 9.2e18L) return edk_fail(EDK_ERR_UNSUPPORTED, "base^L exceeds Int64 contents");
  if (b->L % b->A) return edk_fail(EDK_ERR_ARG, "Length of unit-cell incompatible with L");
  b->len = b->L / b->A;
  b->kp = 0;
  b->P = 1;
  b->ngen = 0;
  b->order = 1;
  int kind = b->kind;
  if (kind == EDK_ABELIAN) {  // basis(...) front door, AbelianBasis.jl:771-847
    int L = b->L;
    std::vector<int32_t> perm(L);
    if (spec->has_k) {
      for (int i = 1; i <= L; i++) perm[i - 1] = ((i - b->A - 1) % L + L) % L + 1;
      if (int rc = add_generator(b, L / b->A, spec->k, perm.data(), nullptr)) return rc;
    }
    if (spec->has_p) {
      for (int i = 1; i <= L; i++) perm[i - 1] = L - i + 1;
      if (int rc = add_generator(b, 2, spec->p == -1 ? 1 : 0, perm.data(), nullptr)) return rc;
    }
    if (spec->has_z) {
      std::vector<uint8_t> inv(L, 1);
      for (int i = 1; i <= L; i++) perm[i - 1] = i;
      if (int rc = add_generator(b, 2, spec->z == -1 ? 1 : 0, perm.data(), inv.data())) return rc;
    }
    for (int c = 0; c < spec->ncustom; c++) {
      const int32_t *pc = spec->perms + (size_t)c * L;
      const uint8_t *ic = spec->invs ? spec->invs + (size_t)c * L : nullptr;
      for (int i = 0; i < L; i++)
        if (pc[i] < 1 || pc[i] > L) return edk_fail(EDK_ERR_ARG, "permutation entries must be in 1:L");
      int ord = compute_period(pc, ic, L);
      if (ord < 1) return edk_fail(EDK_ERR_ARG, "Period computation failed");
      if (int rc = add_generator(b, ord, spec->ks[c], pc, ic)) return rc;
    }
    if (b->ngen == 0) {
      kind = b->kind = spec->N < 0 ? EDK_TENSOR : EDK_PROJECTED;
    } else {
      for (int g = 0; g < b->ngen; g++) b->order *= b->h_gdesc[EDK_GDESC * g + 1];
      if (b->order > 4096) return edk_fail(EDK_ERR_UNSUPPORTED, "symmetry group order above 4096");
      b->h_Rtab.assign(b->order, 0.0);
      for (int i = 1; i <= b->order; i++)  // AbelianBasis.jl:524-527
        if (b->order % i == 0) b->h_Rtab[i - 1] = std::sqrt((double)(b->order * i));
    }
  }
  if (kind == EDK_TRANSLATIONAL) {
    int len = b->len;
    b->kp = ((-spec->k) % len + len) % len;  // TranslationalBasis.jl:247
    b->h_Rtab.resize(len);
    for (int n = 1; n <= len; n++) b->h_Rtab[n - 1] = (double)len / std::sqrt((double)n);  // :250
    b->h_ph.resize(len);
    for (int m = 0; m < len; m++) {  // phase_factor :277-285
      if (b->kp == 0) b->h_ph[m] = make_double2(1.0, 0.0);
      else if (2 * b->kp == len) b->h_ph[m] = make_double2((m % 2 == 0) ? 1.0 : -1.0, 0.0);
      else {
        double th = ((-2.0 * M_PI / len) * b->kp) * m;
        b->h_ph[m] = make_double2(cos(th), sin(th));
      }
    }
    b->order = b->L;
  } else if (kind == EDK_TRANSLATION_PARITY) {
    int len = b->len;
    int km = ((spec->k % len) + len) % len;
    if (km == 0) b->kp = 1;  // TranslationalParityBasis.jl:147-153
    else if ((2 * km) % len == 0) b->kp = -1;
    else return edk_fail(EDK_ERR_ARG, "Momentum incompatible with parity.");
    if (spec->p != 1 && spec->p != -1) return edk_fail(EDK_ERR_ARG, "Invalid parity");
    b->P = spec->p;
    b->h_Rtab.resize(2 * len);
    for (int n = 1; n <= len; n++) {  // N2 = 2len/sqrt(i), N1 = N2/sqrt(2)  (:158-159)
      double n2 = (double)(2 * len) / std::sqrt((double)n);
      b->h_Rtab[len + n - 1] = n2;
      b->h_Rtab[n - 1] = n2 / std::sqrt(2.0);
    }
    b->order = 2 * b->L;
  } else if (kind == EDK_PARITY || kind == EDK_FLIP || kind == EDK_PARITY_FLIP || kind == EDK_TRANSLATION_FLIP) {
    if (spec->p != 1 && spec->p != -1) return edk_fail(EDK_ERR_ARG, "Invalid parity");
    b->P = spec->p;
    if (kind == EDK_PARITY_FLIP) {
      if (spec->z != 1 && spec->z != -1) return edk_fail(EDK_ERR_ARG, "Invalid parity");
      b->Z = spec->z;
    }
    // the flip maps charge N to L(base-1) - N (FlipBasis.jl:77, ParityFlipBasis.jl:92,
    // TranslationalFlipBasis.jl:129)
    if (kind != EDK_PARITY && spec->N >= 0 && 2 * spec->N != b->L * (b->base - 1))
      return edk_fail(EDK_ERR_ARG, "N not compatible with spin flip.");
    if (kind == EDK_PARITY || kind == EDK_FLIP) {
      b->h_Rtab = {std::sqrt(2.0), 2.0};  // ParityJudge / FlipJudge C = sqrt(2); 2.0 for a self-image
      b->order = 2;
    } else if (kind == EDK_PARITY_FLIP) {
      b->h_Rtab = {2.0, 2.0 * std::sqrt(2.0), 0.0, 4.0};  // ParityFlipBasis.jl:96
      b->order = 4;
    } else {
      const int len = b->len;
      b->kp = ((-spec->k) % len + len) % len;  // TranslationalFlipBasis.jl:136
      b->h_Rtab.resize(2 * len);
      for (int n = 1; n <= len; n++) {  // N2 = 2len/sqrt(i), N1 = N2/sqrt(2)  (:140-141)
        const double n2 = (double)(2 * len) / std::sqrt((double)n);
        b->h_Rtab[len + n - 1] = n2;
        b->h_Rtab[n - 1] = n2 / std::sqrt(2.0);
      }
      b->h_ph.resize(len);
      for (int m = 0; m < len; m++) {  // phase_factor(k, len), TranslationalBasis.jl:277-285
        if (b->kp == 0) b->h_ph[m] = make_double2(1.0, 0.0);
        else if (2 * b->kp == len) b->h_ph[m] = make_double2((m % 2 == 0) ? 1.0 : -1.0, 0.0);
        else {
          const double th = ((-2.0 * M_PI / len) * b->kp) * m;
          b->h_ph[m] = make_double2(cos(th), sin(th));
        }
      }
      b->order = 2 * b->L;  // TranslationalFlipBasis.jl:222
    }
  } else if (kind == EDK_PROJECTED) {
    if (spec->N < 0) return edk_fail(EDK_ERR_UNSUPPORTED, "ProjectedBasis needs N (a predicate f is applied on the host)");
  } else if (kind != EDK_TENSOR && kind != EDK_ABELIAN) {
    return edk_fail(EDK_ERR_ARG, "unknown basis kind");
  }
  if (b->h_Rtab.empty()) b->h_Rtab.push_back(1.0);
  if (b->h_ph.empty()) b->h_ph.push_back(make_double2(1.0, 0.0));
  if (b->h_gdesc.empty()) b->h_gdesc.assign(EDK_GDESC, 0);
  if (b->h_gdata.empty()) b->h_gdata.push_back(0);
  if (b->h_gshift.empty()) b->h_gshift.push_back(0);
  EDK_HIP(hipSetDevice(ctx->device));
  if (int rc = upload_vec(ctx, b->h_Rtab, &b->d_Rtab)) return rc;
  if (int rc = upload_vec(ctx, b->h_ph, &b->d_ph)) return rc;
  if (int rc = upload_vec(ctx, b->h_gdesc, &b->d_gdesc)) return rc;
  if (int rc = upload_vec(ctx, b->h_gdata, &b->d_gdata)) return rc;
  if (int rc = upload_vec(ctx, b->h_gshift, &b->d_gshift)) return rc;
  std::vector<uint2> ptab;
  build_prune_table(b, ptab);
  if (!ptab.empty())
    if (int rc = upload_vec(ctx, ptab, &b->d_ptab)) return rc;
  return EDK_OK;
}

static int build_directory(edk_basis *b) {
  edk_ctx *ctx = b->ctx;
  if (b->kind == EDK_TENSOR) return EDK_OK;
  b->maxkey = 0;
  if (b->D > 0)
    EDK_HIP(hipMemcpyAsync(&b->maxkey, b->d_keys + (b->D - 1), 8, hipMemcpyDeviceToHost, ctx->stream));
  EDK_HIP(hipStreamSynchronize(ctx->stream));
  // about one directory slot per representative, at most 2^28 slots
  int64_t target = std::max<int64_t>(1024, std::min<int64_t>(b->D, 1ll << 28));
  int sh = 0;
  while (sh < 63 && (int64_t)(b->maxkey >> sh) + 1 > target) sh++;
  b->dshift = sh;
  b->ndir = (int64_t)(b->maxkey >> sh) + 1;
  EDK_HIP(hipMalloc((void **)&b->d_dir, (b->ndir + 1) * sizeof(uint32_t)));
  int64_t nthr = b->ndir + 1;
  k_build_dir<<<(unsigned)((nthr + 255) / 256), 256, 0, ctx->stream>>>(b->d_keys, b->D, sh, b->ndir, b->d_dir);
  EDK_HIP(hipGetLastError());
  // halving steps of a lower-bound search over the fullest slot: floor(log2 n) + 1
  unsigned int *d_gap = nullptr, gap = 0;
  EDK_HIP(hipMalloc((void **)&d_gap, sizeof(unsigned int)));
  int rc = EDK_OK;
  if (hipMemsetAsync(d_gap, 0, sizeof(unsigned int), ctx->stream) != hipSuccess) rc = edk_fail(EDK_ERR_HIP, "directory");
  if (!rc) {
    k_dir_maxgap<<<(unsigned)((b->ndir + 255) / 256), 256, 0, ctx->stream>>>(b->d_dir, b->ndir, d_gap);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&gap, d_gap, sizeof(unsigned int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = edk_fail(EDK_ERR_HIP, "directory");
  }
  (void)hipFree(d_gap);
  if (rc) return rc;
  int np = 0;
  while (gap >> np) np++;
  b->nprobe = np;
  return EDK_OK;
}

static int run_enumeration(edk_basis *b) {
  edk_ctx *ctx = b->ctx;
  const edk_basis_spec *spec = &b->spec;
  binom_init();
  EnumArgs A;
  A.C = b->canon();
  A.L = b->L;
  A.base = b->base;
  A.w = b->w;
  int num = spec->N >= 0 ? b->L * (b->base - 1) - spec->N : -1;  // sum(dgt) == L(B-1) - N
  if (b->kind == EDK_ABELIAN) {
    if (b->base == 2 && spec->N >= 0 && spec->N <= b->L) {
      A.mode = 0;  // Gosper path: N set bits, predicate dropped (AbelianBasis.jl:530-540)
      A.pop = spec->N;
    } else {
      A.mode = 1;
      A.dsum = num;
    }
  } else if (num >= 0 && b->base == 2) {
    A.mode = 0;
    A.pop = num;
  } else {
    A.mode = 1;
    A.dsum = num;
  }
  if (A.mode == 0) {
    if (A.pop < 0 || A.pop > b->L) {
      b->D = 0;
    }
    A.total = (A.pop < 0 || A.pop > b->L) ? 0 : g_binom[b->L * (EDK_MAXL + 1) + A.pop];
    A.dsum = -1;
  } else {
    A.total = b->pw[b->L];
    A.pop = 0;
  }
  uint64_t *d_binom = nullptr, *d_cnt = nullptr, *d_off = nullptr;
  EDK_HIP(hipMalloc((void **)&d_binom, sizeof(g_binom)));
  EDK_HIP(hipMemcpyAsync(d_binom, g_binom, sizeof(g_binom), hipMemcpyHostToDevice, ctx->stream));
  A.binom = d_binom;
  A.pw = nullptr;
  uint64_t want_threads = (uint64_t)ctx->num_cus * 2048;
  uint64_t per = (A.total + want_threads - 1) / std::max<uint64_t>(want_threads, 1);
  per = std::max<uint64_t>(1, std::min<uint64_t>(per, 256));
  A.per_thread = per;
  uint64_t nthreads = (A.total + per - 1) / per;
  uint64_t nblocks = std::max<uint64_t>(1, (nthreads + ENUM_TPB - 1) / ENUM_TPB);
  EDK_HIP(hipMalloc((void **)&d_cnt, nblocks * sizeof(uint64_t)));
  EDK_HIP(hipMalloc((void **)&d_off, (nblocks + 1) * sizeof(uint64_t)));
  k_enum<false><<<(unsigned)nblocks, ENUM_TPB, 0, ctx->stream>>>(A, d_cnt, nullptr, nullptr, nullptr);
  EDK_HIP(hipGetLastError());
  k_scan_counts<<<1, 1024, 0, ctx->stream>>>(d_cnt, (int64_t)nblocks, d_off);
  EDK_HIP(hipGetLastError());
  uint64_t D = 0;
  EDK_HIP(hipMemcpyAsync(&D, d_off + nblocks, 8, hipMemcpyDeviceToHost, ctx->stream));
  EDK_HIP(hipStreamSynchronize(ctx->stream));
  b->D = (int64_t)D;
  hipError_t e1 = hipMalloc((void **)&b->d_keys, std::max<uint64_t>(D, 1) * sizeof(uint64_t));
  hipError_t e2 = hipMalloc((void **)&b->d_R, std::max<uint64_t>(D, 1) * sizeof(double));
  if (e1 != hipSuccess || e2 != hipSuccess) {
    hipFree(d_binom);
    hipFree(d_cnt);
    hipFree(d_off);
    return edk_fail(EDK_ERR_OOM, "out of device memory for the representative table");
  }
  k_enum<true><<<(unsigned)nblocks, ENUM_TPB, 0, ctx->stream>>>(A, d_cnt, d_off, b->d_keys, b->d_R);
  EDK_HIP(hipGetLastError());
  EDK_HIP(hipStreamSynchronize(ctx->stream));
  hipFree(d_binom);
  hipFree(d_cnt);
  hipFree(d_off);
  return build_directory(b);
}

extern "C" int edk_basis_build(edk_ctx *ctx, const edk_basis_spec *spec, edk_basis **out) {
  if (!ctx || !spec || !out) return edk_fail(EDK_ERR_ARG, "edk_basis_build: null argument");
  edk_basis *b = new edk_basis();
  int rc = setup_basis(ctx, spec, b);
  if (rc == EDK_OK) {
    if (b->kind == EDK_TENSOR) {
      b->D = (int64_t)b->pw[b->L];
    } else {
      rc = run_enumeration(b);
    }
  }
  if (rc != EDK_OK) {
    free_basis_mem(b);
    delete b;
    return rc;
  }
  *out = b;
  return EDK_OK;
}

extern "C" int edk_basis_upload(edk_ctx *ctx, const edk_basis_spec *spec, const int64_t *I1,
                                const double *R, int64_t D, edk_basis **out) {
  if (!ctx || !spec || !out || D < 0 || (D > 0 && !I1)) return edk_fail(EDK_ERR_ARG, "edk_basis_upload: bad argument");
  edk_basis *b = new edk_basis();
  int rc = setup_basis(ctx, spec, b);
  if (rc == EDK_OK && b->kind == EDK_TENSOR) rc = edk_fail(EDK_ERR_ARG, "TensorBasis has no table to upload");
  if (rc == EDK_OK) {
    std::vector<uint64_t> keys(D);
    std::vector<double> Rv(D, 1.0);
    uint64_t fm = (1ull << b->w) - 1;
    (void)fm;
    for (int64_t i = 0; i < D; i++) {
      if (I1[i] < 1 || (uint64_t)(I1[i] - 1) >= b->pw[b->L]) {
        rc = edk_fail(EDK_ERR_ARG, "representative outside 1:base^L");
        break;
      }
      if (i && I1[i] <= I1[i - 1]) {
        rc = edk_fail(EDK_ERR_ARG, "representatives must be strictly increasing");
        break;
      }
      uint64_t c = (uint64_t)(I1[i] - 1), s = 0;
      for (int q = 0; q < b->L; q++) {
        s |= (c % (uint64_t)b->base) << (q * b->w);
        c /= (uint64_t)b->base;
      }
      keys[i] = s;
      if (R) Rv[i] = R[i];
    }
    b->D = D;
    if (rc == EDK_OK) rc = upload_vec(ctx, keys, &b->d_keys);
    if (rc == EDK_OK) rc = upload_vec(ctx, Rv, &b->d_R);
    if (rc == EDK_OK) rc = build_directory(b);
  }
  if (rc != EDK_OK) {
    free_basis_mem(b);
    delete b;
    return rc;
  }
  b->uploaded = true;
  *out = b;
  return EDK_OK;
}

extern "C" int64_t edk_basis_dim(const edk_basis *b) { return b ? b->D : -1; }
extern "C" int edk_basis_order(const edk_basis *b) { return b ? b->order : -1; }

extern "C" int edk_basis_download(edk_basis *b, int64_t *I1, double *R) {
  if (!b) return edk_fail(EDK_ERR_ARG, "null basis");
  edk_ctx *ctx = b->ctx;
  EDK_HIP(hipSetDevice(ctx->device));
  if (b->kind == EDK_TENSOR) {
    for (int64_t i = 0; i < b->D; i++) {
      if (I1) I1[i] = i + 1;
      if (R) R[i] = 1.0;
    }
    return EDK_OK;
  }
  if (b->D == 0) return EDK_OK;
  if (I1) {
    int64_t *d_c = nullptr;
    EDK_HIP(hipMalloc((void **)&d_c, b->D * sizeof(int64_t)));
    k_packed_to_content<<<(unsigned)((b->D + 255) / 256), 256, 0, ctx->stream>>>(b->d_keys, b->D, b->L, b->base,
                                                                               b->w, d_c);
    EDK_HIP(hipGetLastError());
    EDK_HIP(hipMemcpyAsync(I1, d_c, b->D * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    EDK_HIP(hipStreamSynchronize(ctx->stream));
    hipFree(d_c);
  }
  if (R) {
    EDK_HIP(hipMemcpyAsync(R, b->d_R, b->D * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    EDK_HIP(hipStreamSynchronize(ctx->stream));
    if (b->kind == EDK_PROJECTED)
      for (int64_t i = 0; i < b->D; i++) R[i] = 1.0;
  }
  return EDK_OK;
}

extern "C" int edk_basis_index(edk_basis *b, const int64_t *contents, int64_t n, int64_t *pos,
                               double *cre, double *cim) {
  if (!b || (n > 0 && (!contents || !pos || !cre || !cim))) return edk_fail(EDK_ERR_ARG, "edk_basis_index: bad argument");
  if (n == 0) return EDK_OK;
  edk_ctx *ctx = b->ctx;
  for (int64_t q = 0; q < n; q++)
    if (contents[q] < 0 || (uint64_t)contents[q] >= b->pw[b->L])
      return edk_fail(EDK_ERR_ARG, "content outside 0:base^L-1");
  EDK_HIP(hipSetDevice(ctx->device));
  int64_t *d_in = nullptr, *d_pos = nullptr;
  double *d_re = nullptr, *d_im = nullptr;
  EDK_HIP(hipMalloc((void **)&d_in, n * 8));
  EDK_HIP(hipMalloc((void **)&d_pos, n * 8));
  EDK_HIP(hipMalloc((void **)&d_re, n * 8));
  EDK_HIP(hipMalloc((void **)&d_im, n * 8));
  EDK_HIP(hipMemcpyAsync(d_in, contents, n * 8, hipMemcpyHostToDevice, ctx->stream));
  k_index<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(b->canon(), b->search(), b->d_R, b->kind, b->L,
                                                                b->base, b->w, d_in, n, d_pos, d_re, d_im);
  EDK_HIP(hipGetLastError());
  EDK_HIP(hipMemcpyAsync(pos, d_pos, n * 8, hipMemcpyDeviceToHost, ctx->stream));
  EDK_HIP(hipMemcpyAsync(cre, d_re, n * 8, hipMemcpyDeviceToHost, ctx->stream));
  EDK_HIP(hipMemcpyAsync(cim, d_im, n * 8, hipMemcpyDeviceToHost, ctx->stream));
  EDK_HIP(hipStreamSynchronize(ctx->stream));
  hipFree(d_in);
  hipFree(d_pos);
  hipFree(d_re);
  hipFree(d_im);
  return EDK_OK;
}

extern "C" int edk_basis_destroy(edk_basis *b) {
  if (!b) return EDK_OK;
  if (b->refcount > 0) return edk_fail(EDK_ERR_STATE, "basis still referenced by an operator");
  hipSetDevice(b->ctx->device);
  hipStreamSynchronize(b->ctx->stream);
  free_basis_mem(b);
  delete b;
  return EDK_OK;
}
