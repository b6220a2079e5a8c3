// edk_internal.h -- shared host/device definitions of libedkit_hip.
//
// Packed states.  A product state is a uint64 holding one w-bit field per
// site, site 1 (the reference's most significant digit, AbstractBasis.jl:
// 285-298) in the highest field: field of site s = bits [(L-s)*w, (L-s+1)*w).
// w = ceil(log2(base)).  For base 2 the packed word IS the reference's 0-based
// content I[i]-1; for other bases the packed order equals the base-B integer
// order (both are lexicographic in the digits), so a table sorted by content
// is sorted by packed word and every "strict <" of the reference's orbit
// searches picks the same element.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/edkit_hip.h"

#define EDK_MAXL 64
#define EDK_MAXGEN 8
#define EDK_MAXSITES 8

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
void edk_set_error(const std::string &msg);
int edk_fail(int code, const std::string &msg);
// device address of edk_vec_dot_noalloc's scratch array, once per device (device current)
int edk_noalloc_setup(int device);
#define EDK_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return edk_fail(EDK_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct edk_ctx {
  int device;
  hipStream_t stream;
  bool own_stream;
  int num_cus;
  // cached allocation for the Krylov drivers' per-call buffers: blocks come
  // from a per-device cache of size classes and go back to it after this
  // context's stream has finished with them.  A plain hipFree waits for the
  // whole device (serialising contexts that run side by side), and fresh
  // device memory costs ~20-70 GB/s to map.
  bool cache_alloc = false;
  // two pinned slots of EDK_PINNED_SLOT double2 for reduction results, owned
  // by the context (no pinned allocation per call)
  void *pinned = nullptr;
};
#define EDK_PINNED_SLOT 1040

// device buffers of a context's calls (see edk_ctx::cache_alloc)
hipError_t edk_dalloc(edk_ctx *c, void **p, size_t bytes);
void edk_dfree(edk_ctx *c, void *p);
// bytes the block cache of a device holds for reuse
size_t edk_cache_bytes(int device);

// Generator kinds of the Abelian canonicaliser.
// GEN_SHIFTS: a site permutation whose fields move by at most EDK_MAXSHIFTS
// distinct bit displacements, applied as OR_g shift(s & mask_g, d_g)
enum { GEN_ROTL = 0, GEN_REV = 1, GEN_IDENT = 2, GEN_BENES = 3, GEN_SHIFTS = 4 };
#define EDK_MAXSHIFTS 4

// Uniform description of a basis' canonicaliser, passed by value to kernels.
// Tables it points to are small device arrays owned by the basis.
struct CanonDev {
  int kind, L, w, A, len, base;
  uint64_t mask;        // low L*w bits
  uint64_t lowfields;   // one bit at the bottom of every field (w == 2 flips)
  int kp, P;            // Translational: k' ; TP: K (+-1), P (+-1); flip bases: P = flip or reflection eigenvalue
  int Z;                // ParityFlipBasis: spin-flip eigenvalue (+-1)
  int ngen, order;
  const double *Rtab;   // norms by orbit data (see host setup)
  const double2 *ph;    // phase tables
  const int *gdesc;     // per generator: type, order, shift/param, ph offset, layer offset, nlayers
  const uint64_t *gdata;// Benes layer masks, inversion masks
  const int *gshift;    // Benes layer shifts
  // pruned Abelian walk (null: the full walk): generator 0 maps the top
  // pbits of the word onto themselves; per value v of those bits, the least
  // top block over generator 0's powers and the mask of the powers reaching it
  const uint2 *ptab;
  int pbits;
};

struct SearchDev {
  const uint64_t *keys;  // sorted packed representatives
  int64_t D;
  const uint32_t *dir;   // dir[b] = first index with key >= (b << dshift), b in [0, ndir]
  int dshift;
  int64_t ndir;
  int nprobe;            // halving steps that resolve the fullest directory slot
};

struct edk_basis {
  edk_ctx *ctx;
  edk_basis_spec spec;
  int kind, L, base, w, A, len;
  int kp, P, Z = 1;
  int64_t D;
  uint64_t maxkey;
  uint64_t *d_keys = nullptr;
  double *d_R = nullptr;
  uint32_t *d_dir = nullptr;
  int dshift = 0;
  int64_t ndir = 0;
  int nprobe = 0;
  // canonicaliser tables
  double *d_Rtab = nullptr;
  double2 *d_ph = nullptr;
  int *d_gdesc = nullptr;
  uint64_t *d_gdata = nullptr;
  int *d_gshift = nullptr;
  int ngen = 0, order = 1;
  std::vector<double> h_Rtab;
  std::vector<double2> h_ph;
  std::vector<int> h_gdesc;
  std::vector<uint64_t> h_gdata;
  std::vector<int> h_gshift;
  uint2 *d_ptab = nullptr;   // pruned Abelian walk table (CanonDev::ptab)
  int pbits = 0;
  std::vector<int32_t> h_perm0;  // generator 0's permutation and inversion flags
  std::vector<uint8_t> h_inv0;
  std::vector<uint64_t> pw;  // base^i
  int refcount = 0;
  bool uploaded = false;  // table came from the caller: gather R, never infer it
  CanonDev canon() const;
  SearchDev search() const;
};

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t edk_rotr(uint64_t s, int sh, int tot, uint64_t mask) {
  // rotate the low `tot` bits right by sh (0 < sh < tot)
  return ((s >> sh) | (s << (tot - sh))) & mask;
}

// reverse the order of the L w-bit fields
__device__ __forceinline__ uint64_t edk_revfields(uint64_t s, int L, int w) {
  if (w == 1) return __brevll(s) >> (64 - L);
  if (w == 2) {
    uint64_t t = __brevll(s) >> (64 - 2 * L);
    return ((t & 0x5555555555555555ull) << 1) | ((t >> 1) & 0x5555555555555555ull);
  }
  uint64_t r = 0, fm = (1ull << w) - 1;
  for (int i = 0; i < L; i++) {
    r = (r << w) | (s & fm);
    s >>= w;
  }
  return r;
}

__device__ __forceinline__ int64_t edk_search(const SearchDev &S, uint64_t key) {
  uint64_t b = key >> S.dshift;
  if (b >= (uint64_t)S.ndir) return -1;
  int64_t lo = S.dir[b], hi = S.dir[b + 1];
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    uint64_t v = S.keys[mid];
    if (v < key) lo = mid + 1;
    else hi = mid;
  }
  if (lo < S.D && S.keys[lo] == key) return lo;
  return -1;
}

// Abelian generator application on a packed word.
__device__ __forceinline__ uint64_t edk_apply_gen(const CanonDev &C, int g, uint64_t s) {
  const int *d = C.gdesc + 8 * g;
  int type = d[0];
  int tot = C.L * C.w;
  if (type == GEN_ROTL) {
    int sh = d[2];
    if (sh) s = ((s << sh) | (s >> (tot - sh))) & C.mask;
  } else if (type == GEN_REV) {
    s = edk_revfields(s, C.L, C.w);
  } else if (type == GEN_BENES) {
    int off = d[4], n = d[5];
    for (int q = 0; q < n; q++) {
      uint64_t m = C.gdata[off + q];
      int sh = C.gshift[off + q];
      uint64_t dd = ((s >> sh) ^ s) & m;
      s = s ^ dd ^ (dd << sh);
    }
  } else if (type == GEN_SHIFTS) {
    // fields grouped by bit displacement (positive: towards the MSB)
    int off = d[4], n = d[5];
    uint64_t r = 0;
    for (int q = 0; q < n; q++) {
      const uint64_t m = s & C.gdata[off + q];
      const int sh = C.gshift[off + q];
      r |= sh >= 0 ? (m << sh) : (m >> -sh);
    }
    s = r;
  }
  // spin inversion: gdata[ngen-indexed inversion slot]
  // spin inversion d -> base-1-d on the flagged sites (apply_inv!,
  // AbelianBasis.jl:268-283).  For base 2 and 4 that is an XOR of the whole
  // field; for base 3 (00 <-> 10, 01 fixed) the mask holds the fields' high
  // bits and only fields whose low bit is clear flip.
  uint64_t inv = C.gdata[d[3]];
  if (inv) {
    if (C.base != 3) s ^= inv;
    else s ^= inv & ~((s & C.lowfields) << 1);
  }
  return s;
}

// Result of canonicalising a configuration.
struct CanonOut {
  uint64_t rep;
  int ok;        // 0: the representative cannot be in the table
  double R;      // norm of the representative (same FP value as the table)
  double2 phase; // the reference's coefficient phase (multiplies R)
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// TranslationalBasis index() (TranslationalBasis.jl:310-341) plus the orbit
// period, from which the norm len/sqrt(period) follows (the judge, :94-114).
__device__ __forceinline__ CanonOut edk_canon_trans(const CanonDev &C, uint64_t s) {
  int tot = C.L * C.w, sh = C.A * C.w;
  uint64_t st = s, best = s;
  int M = 0, period = C.len;
  for (int i = 1; i < C.len; i++) {
    st = edk_rotr(st, sh, tot, C.mask);
    if (st == s) {
      period = i;
      break;
    }
    if (st < best) {
      best = st;
      M = i;
    }
  }
  CanonOut o;
  o.rep = best;
  o.ok = ((int64_t)period * C.kp) % C.len == 0;  // check_momentum (:80)
  o.R = C.Rtab[period - 1];
  o.phase = C.ph[M];
  return o;
}

// TranslationParityBasis index() (TranslationalParityBasis.jl:192-234).
// The orbit is reflection-symmetric iff the reflected minimum equals the
// translated minimum; the norm is then N2[period], else N1[period] (:158-159).
__device__ __forceinline__ CanonOut edk_canon_tp(const CanonDev &C, uint64_t s) {
  int tot = C.L * C.w, sh = C.A * C.w;
  uint64_t st = s;
  uint64_t Iam = s, Ibm = edk_revfields(s, C.L, C.w);
  int Ri = 0, Mi = 0, period = C.len;
  for (int i = 1; i < C.len; i++) {
    st = edk_rotr(st, sh, tot, C.mask);
    if (st == s) {
      period = i;
      break;
    }
    if (st < Iam) {
      Iam = st;
      Ri = i;
    }
    uint64_t rb = edk_revfields(st, C.L, C.w);
    if (rb < Ibm) {
      Ibm = rb;
      Mi = i;
    }
  }
  CanonOut o;
  bool qp = (Iam == Ibm);
  double sgn;
  if (Ibm < Iam) {
    o.rep = Ibm;
    sgn = (double)C.P * ((C.kp == 1 || (Mi & 1) == 0) ? 1.0 : -1.0);
  } else {
    o.rep = Iam;
    sgn = (C.kp == 1 || (Ri & 1) == 0) ? 1.0 : -1.0;
  }
  o.ok = 1;
  o.R = C.Rtab[(qp ? C.len : 0) + period - 1];
  o.phase = make_double2(sgn, 0.0);
  return o;
}

// AbelianBasis index() (AbelianBasis.jl:669-693 via shift_canonical_int
// :428-445): walk all order(G) elements in odometer order, keep the first
// strict minimum.  The stabiliser count gives the norm sqrt(Ng*n) (:524-527).
// Generator descriptor (EDK_GDESC ints each): type, order, rotation bits,
// inversion-mask slot, Benes layer offset, layer count, phase-table offset.
#define EDK_GDESC 8

// A generator held in registers for the walk (every form but Benes).
struct GenReg {
  int type, order, rot, n;
  uint64_t inv;
  uint64_t m[EDK_MAXSHIFTS];
  int sh[EDK_MAXSHIFTS];
};

__device__ __forceinline__ GenReg edk_load_gen(const CanonDev &C, int g) {
  const int *d = C.gdesc + EDK_GDESC * g;
  GenReg r;
  r.type = d[0];
  r.order = d[1];
  r.rot = d[2];
  r.inv = C.gdata[d[3]];
  r.n = r.type == GEN_SHIFTS ? d[5] : 0;
#pragma unroll
  for (int q = 0; q < EDK_MAXSHIFTS; q++) {
    r.m[q] = q < r.n ? C.gdata[d[4] + q] : 0ull;
    r.sh[q] = q < r.n ? C.gshift[d[4] + q] : 0;
  }
  return r;
}

// edk_apply_gen for a register-resident generator, on a word type W that
// holds the L w-bit fields (uint32_t when L w <= 32: one instruction per
// logical operation instead of two)
template <typename W>
__device__ __forceinline__ W edk_apply_genreg(const CanonDev &C, const GenReg &r, W s) {
  if (r.type == GEN_ROTL) {
    const int tot = C.L * C.w;
    if (r.rot) s = ((s << r.rot) | (s >> (tot - r.rot))) & (W)C.mask;
  } else if (r.type == GEN_REV) {
    s = (W)edk_revfields((uint64_t)s, C.L, C.w);
  } else if (r.type == GEN_SHIFTS) {
    W o = 0;
#pragma unroll
    for (int q = 0; q < EDK_MAXSHIFTS; q++) {
      if (q < r.n) {
        const W m = s & (W)r.m[q];
        o |= r.sh[q] >= 0 ? (W)(m << r.sh[q]) : (W)(m >> -r.sh[q]);
      }
    }
    s = o;
  }
  if (r.inv) {
    if (C.base != 3) s ^= (W)r.inv;
    else s ^= (W)r.inv & (W) ~((s & (W)C.lowfields) << 1);
  }
  return s;
}

// The walk of edk_canon_abelian with up to three generators, none of them a
// Benes network, held in registers: nested loops over the exponents
// (generator 0 fastest, the odometer order), so the loop control is uniform
// and each element costs one generator application; the first strict
// minimum is kept as its element number e = a0 + n0 (a1 + n1 a2).
template <typename W>
__device__ __forceinline__ CanonOut edk_canon_abelian3(const CanonDev &C, uint64_t s64) {
  const GenReg g0 = edk_load_gen(C, 0);
  const GenReg g1 = C.ngen > 1 ? edk_load_gen(C, 1) : g0;
  const GenReg g2 = C.ngen > 2 ? edk_load_gen(C, 2) : g0;
  const int n0 = g0.order, n1 = C.ngen > 1 ? g1.order : 1, n2 = C.ngen > 2 ? g2.order : 1;
  const W s = (W)s64;
  W best = s;
  int be = 0, nstab = 0, e = 0;
  W c2 = s;
  for (int a2 = 0; a2 < n2; a2++) {
    W c1 = c2;
    for (int a1 = 0; a1 < n1; a1++) {
      W c0 = c1;
      for (int a0 = 0; a0 < n0; a0++, e++) {
        nstab += c0 == s;
        if (c0 < best) {
          best = c0;
          be = e;
        }
        if (a0 + 1 < n0) c0 = edk_apply_genreg<W>(C, g0, c0);
      }
      if (a1 + 1 < n1) c1 = edk_apply_genreg<W>(C, g1, c1);
    }
    if (a2 + 1 < n2) c2 = edk_apply_genreg<W>(C, g2, c2);
  }
  const int m0 = be % n0 + 1, m1 = (be / n0) % n1 + 1, m2 = be / (n0 * n1) + 1;
  CanonOut o;
  o.rep = (uint64_t)best;
  o.ok = 1;
  o.R = C.Rtab[nstab - 1];
  double2 acc = C.ph[C.gdesc[6] + m0 - 1];
  if (C.ngen > 1) acc = cmul(acc, C.ph[C.gdesc[EDK_GDESC + 6] + m1 - 1]);
  if (C.ngen > 2) acc = cmul(acc, C.ph[C.gdesc[2 * EDK_GDESC + 6] + m2 - 1]);
  o.phase = acc;
  return o;
}

// The same walk pruned by the top block (CanonDev::ptab; two or three
// generators, none of them a Benes network).  Generator 0 maps the top pbits
// onto themselves, so an image g0^a0 g'(s) has the top block
// g0^a0(top(g'(s))) and the least top block over a0 is a table entry.  Pass 1
// finds the least top block V over the outer elements g' = g1^a1 g2^a2; only
// images with top block V can be the minimum or equal it, and pass 2 compares
// just those (the table's mask of a0) in odometer order, keeping the first
// strict minimum.  The elements mapping s onto its representative are a coset
// of the stabiliser, so counting the images equal to the minimum gives the
// stabiliser order the norm needs.
// With at most EDK_PRUNE_OUT outer elements they are held in registers and
// their table entries loaded together.
#define EDK_PRUNE_OUT 8
template <typename W>
__device__ __forceinline__ CanonOut edk_canon_abelian_pruned(const CanonDev &C, uint64_t s64) {
  const GenReg g0 = edk_load_gen(C, 0), g1 = edk_load_gen(C, 1);
  const GenReg g2 = C.ngen > 2 ? edk_load_gen(C, 2) : g1;
  const int n0 = g0.order, n1 = g1.order, n2 = C.ngen > 2 ? g2.order : 1;
  const int sh = C.L * C.w - C.pbits;
  const W s = (W)s64;
  W best = s;
  int be = 0, cnt = 0;
  if (n1 * n2 <= EDK_PRUNE_OUT) {
    // outer elements in odometer order (generator 1 fastest), then their
    // table entries, all loads in flight together
    W co[EDK_PRUNE_OUT];
    W c1 = s, c2 = s;
    int a1 = 0;
#pragma unroll
    for (int i = 0; i < EDK_PRUNE_OUT; i++) {
      co[i] = c1;
      if (++a1 < n1) {
        c1 = edk_apply_genreg<W>(C, g1, c1);
      } else {
        a1 = 0;
        c2 = edk_apply_genreg<W>(C, g2, c2);
        c1 = c2;
      }
    }
    uint2 te[EDK_PRUNE_OUT];
#pragma unroll
    for (int i = 0; i < EDK_PRUNE_OUT; i++) te[i] = C.ptab[i < n1 * n2 ? (uint32_t)(co[i] >> sh) : 0u];
    uint32_t vmin = 0xffffffffu;
#pragma unroll
    for (int i = 0; i < EDK_PRUNE_OUT; i++) vmin = i < n1 * n2 ? min(vmin, te[i].x) : vmin;
#pragma unroll
    for (int i = 0; i < EDK_PRUNE_OUT; i++) {
      uint32_t mask = (i < n1 * n2 && te[i].x == vmin) ? te[i].y : 0u;
      W c0 = co[i];
      int a0 = 0;
      while (mask) {
        const int nx = __ffs(mask) - 1;
        mask &= mask - 1;
        for (; a0 < nx; a0++) c0 = edk_apply_genreg<W>(C, g0, c0);
        const int e = a0 + n0 * i;
        if (cnt == 0 || c0 < best) {
          best = c0;
          be = e;
          cnt = 1;
        } else if (c0 == best) {
          cnt++;
        }
      }
    }
  } else {
    uint32_t vmin = 0xffffffffu;
    W c2 = s;
    for (int a2 = 0; a2 < n2; a2++) {
      W c1 = c2;
      for (int a1 = 0; a1 < n1; a1++) {
        vmin = min(vmin, C.ptab[(uint32_t)(c1 >> sh)].x);
        if (a1 + 1 < n1) c1 = edk_apply_genreg<W>(C, g1, c1);
      }
      if (a2 + 1 < n2) c2 = edk_apply_genreg<W>(C, g2, c2);
    }
    c2 = s;
    for (int a2 = 0; a2 < n2; a2++) {
      W c1 = c2;
      for (int a1 = 0; a1 < n1; a1++) {
        const uint2 t = C.ptab[(uint32_t)(c1 >> sh)];
        if (t.x == vmin) {
          uint32_t mask = t.y;
          W c0 = c1;
          int a0 = 0;
          while (mask) {
            const int nx = __ffs(mask) - 1;
            mask &= mask - 1;
            for (; a0 < nx; a0++) c0 = edk_apply_genreg<W>(C, g0, c0);
            const int e = a0 + n0 * (a1 + n1 * a2);
            if (cnt == 0 || c0 < best) {
              best = c0;
              be = e;
              cnt = 1;
            } else if (c0 == best) {
              cnt++;
            }
          }
        }
        if (a1 + 1 < n1) c1 = edk_apply_genreg<W>(C, g1, c1);
      }
      if (a2 + 1 < n2) c2 = edk_apply_genreg<W>(C, g2, c2);
    }
  }
  const int m0 = be % n0 + 1, m1 = (be / n0) % n1 + 1, m2 = be / (n0 * n1) + 1;
  CanonOut o;
  o.rep = (uint64_t)best;
  o.ok = 1;
  o.R = C.Rtab[cnt - 1];
  double2 acc = C.ph[C.gdesc[6] + m0 - 1];
  acc = cmul(acc, C.ph[C.gdesc[EDK_GDESC + 6] + m1 - 1]);
  if (C.ngen > 2) acc = cmul(acc, C.ph[C.gdesc[2 * EDK_GDESC + 6] + m2 - 1]);
  o.phase = acc;
  return o;
}

__device__ __forceinline__ CanonOut edk_canon_abelian(const CanonDev &C, uint64_t s) {
  if (C.ptab) return C.L * C.w <= 32 ? edk_canon_abelian_pruned<uint32_t>(C, s) : edk_canon_abelian_pruned<uint64_t>(C, s);
  if (C.ngen >= 1 && C.ngen <= 3 && C.gdesc[0] != GEN_BENES && (C.ngen < 2 || C.gdesc[EDK_GDESC] != GEN_BENES) &&
      (C.ngen < 3 || C.gdesc[2 * EDK_GDESC] != GEN_BENES))
    return C.L * C.w <= 32 ? edk_canon_abelian3<uint32_t>(C, s) : edk_canon_abelian3<uint64_t>(C, s);
  int os[EDK_MAXGEN], ms[EDK_MAXGEN];
  for (int g = 0; g < C.ngen; g++) os[g] = ms[g] = 1;
  uint64_t best = s, cur = s;
  int nstab = 1;
  for (int step = 1; step < C.order; step++) {
    for (int g = 0; g < C.ngen; g++) {
      cur = edk_apply_gen(C, g, cur);
      if (++os[g] > C.gdesc[EDK_GDESC * g + 1]) os[g] = 1;
      else break;
    }
    if (cur == s) nstab++;
    if (cur < best) {
      best = cur;
      for (int g = 0; g < C.ngen; g++) ms[g] = os[g];
    }
  }
  CanonOut o;
  o.rep = best;
  o.ok = 1;
  o.R = C.Rtab[nstab - 1];
  // phase = prod_g c_g[s_g] (AbelianBasis.jl:225-227), left to right
  double2 acc = C.ph[C.gdesc[6] + ms[0] - 1];
  for (int g = 1; g < C.ngen; g++) acc = cmul(acc, C.ph[C.gdesc[EDK_GDESC * g + 6] + ms[g] - 1]);
  o.phase = acc;
  return o;
}

// spin flip of every field: digit d -> base - 1 - d (_int_spinflip,
// AbstractBasis.jl:434: maxstate - state); no borrows, each field d <= base-1
__device__ __forceinline__ uint64_t edk_flipfields(const CanonDev &C, uint64_t s) {
  return (uint64_t)(C.base - 1) * C.lowfields - s;
}

// ParityBasis index() (ParityBasis.jl:110-121) and FlipBasis index()
// (FlipBasis.jl:105-115): the smaller of the state and its image; the image
// carries the eigenvalue P.  Norm 2 for a self-image (kept only at P = +1), else
// sqrt(2) (judges ParityBasis.jl:44-57, FlipBasis.jl:42-55).
__device__ __forceinline__ CanonOut edk_canon_pair(const CanonDev &C, uint64_t s) {
  const uint64_t b = C.kind == EDK_PARITY ? edk_revfields(s, C.L, C.w) : edk_flipfields(C, s);
  CanonOut o;
  const bool self = (b == s);
  o.rep = b < s ? b : s;
  o.ok = !self || C.P == 1;
  o.R = C.Rtab[self ? 1 : 0];
  o.phase = make_double2(b < s ? (double)C.P : 1.0, 0.0);
  return o;
}

// ParityFlipBasis index() (ParityFlipBasis.jl:122-145): the minimum of the
// state, its reflection, its flip and the flipped reflection; the first of
// them (in that order) equal to the minimum fixes the factor 1, P, Z or PZ.
// The norm follows the stabiliser count, C = [2, 2sqrt(2), 0, 4] (:96),
// and a stabiliser with eigenvalue -1 empties the orbit (double_parity_check
// :47-65).
__device__ __forceinline__ CanonOut edk_canon_parityflip(const CanonDev &C, uint64_t s) {
  const uint64_t Ip = edk_revfields(s, C.L, C.w), Iz = edk_flipfields(C, s), Ipz = edk_flipfields(C, Ip);
  uint64_t r = s < Ip ? s : Ip;
  r = r < Iz ? r : Iz;
  r = r < Ipz ? r : Ipz;
  double f;
  if (r == s) f = 1.0;
  else if (r == Ip) f = (double)C.P;
  else if (r == Iz) f = (double)C.Z;
  else f = (double)(C.P * C.Z);
  int n = 1;
  bool ok = true;
  if (Ip == s) {
    n++;
    ok = ok && C.P == 1;
  }
  if (Iz == s) {
    n++;
    ok = ok && C.Z == 1;
  }
  if (Ipz == s) {
    n++;
    ok = ok && C.P * C.Z == 1;
  }
  CanonOut o;
  o.rep = r;
  o.ok = ok;
  o.R = C.Rtab[n - 1];
  o.phase = make_double2(f, 0.0);
  return o;
}

// TranslationFlipBasis index() (TranslationalFlipBasis.jl:172-215): right
// rotations by the unit cell until the state returns; the first strict
// minimum of the rotations (count R) and of their flips (count M, starting
// from the flipped state itself).  A flipped minimum strictly below the
// translated one gives P * C[M], else C[R].  The norm is N2[period] when some
// flipped rotation equals the state, else N1[period] (:140-141, judge :50-94).
__device__ __forceinline__ CanonOut edk_canon_transflip(const CanonDev &C, uint64_t s) {
  const int tot = C.L * C.w, sh = C.A * C.w;
  uint64_t st = s, Iam = s, Ibm = edk_flipfields(C, s);
  int R = 0, M = 0, period = C.len;
  bool qp = (Ibm == s);
  for (int i = 1; i < C.len; i++) {
    st = edk_rotr(st, sh, tot, C.mask);
    if (st == s) {
      period = i;
      break;
    }
    if (st < Iam) {
      Iam = st;
      R = i;
    }
    const uint64_t fb = edk_flipfields(C, st);
    if (fb < Ibm) {
      Ibm = fb;
      M = i;
    }
    qp = qp || fb == s;
  }
  CanonOut o;
  o.ok = ((int64_t)period * C.kp) % C.len == 0;  // absent otherwise: the judge's momentum test
  o.R = C.Rtab[(qp ? C.len : 0) + period - 1];
  if (Ibm < Iam) {
    o.rep = Ibm;
    const double2 c = C.ph[M];
    o.phase = make_double2(C.P * c.x, C.P * c.y);
  } else {
    o.rep = Iam;
    o.phase = C.ph[R];
  }
  return o;
}

__device__ __forceinline__ CanonOut edk_canon(const CanonDev &C, uint64_t s) {
  if (C.kind == EDK_TRANSLATIONAL) return edk_canon_trans(C, s);
  if (C.kind == EDK_TRANSLATION_PARITY) return edk_canon_tp(C, s);
  if (C.kind == EDK_ABELIAN) return edk_canon_abelian(C, s);
  if (C.kind == EDK_PARITY || C.kind == EDK_FLIP) return edk_canon_pair(C, s);
  if (C.kind == EDK_PARITY_FLIP) return edk_canon_parityflip(C, s);
  if (C.kind == EDK_TRANSLATION_FLIP) return edk_canon_transflip(C, s);
  CanonOut o;
  o.rep = s;
  o.ok = 1;
  o.R = 1.0;
  o.phase = make_double2(1.0, 0.0);
  return o;
}
