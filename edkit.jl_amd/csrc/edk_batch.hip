// edk_batch.hip -- multi-sector matvec batches as one HIP graph, and the
// schedule autotuner's C ABI.
//
// A batch records the matvecs of its items (operator, x, y, block width) once
// under stream capture and replays them with a single hipGraphLaunch.  Block
// products of sectors that take the same specialised block kernel (same
// basis kind, phase mode, length class, block width and accumulate flag) are
// fused into one launch over all their chunks (k_matvec_bal_blk_multi), up to
// 12 sectors a launch; the other items keep a launch each.  Launches are
// spread over `streams` capture branches (fork from the context stream, join
// back), balanced by D * nb, so the graph has that many independent chains
// and small sectors do not wait for each other.  Capture records no memset
// and no allocation: every kernel writes all of its y (accumulate = 0) or
// reads it (accumulate = 1), and the kernels' one-time setup runs before
// capture begins.
#include <algorithm>
#include <map>
#include <tuple>
#include <vector>

#include "edk_fast.h"

struct edk_batch {
  edk_ctx *ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t n = 0;
  int64_t launches = 0;        // kernel launches the graph holds
};

extern "C" int edk_operator_set_variant(edk_op *op, int dtype, int64_t nb, int variant) {
  if (!op) return edk_fail(EDK_ERR_ARG, "null operator");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "bad dtype");
  EDK_HIP(hipSetDevice(op->ctx->device));
  return edk_fast_set_variant(op, dtype, nb, variant);
}

extern "C" int edk_operator_autotune(edk_op *op, int dtype, int64_t nb, int reps, int *variant, double *ms) {
  if (!op) return edk_fail(EDK_ERR_ARG, "null operator");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "bad dtype");
  EDK_HIP(hipSetDevice(op->ctx->device));
  return edk_fast_autotune(op, dtype, nb, reps, variant, ms);
}

static void batch_free(edk_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  delete b;
}

// Fusion key of an item's block product, or an empty key (fused == false)
// when the item keeps its own launch: the specialised block kernel must be
// the one its launch would take (set up, chunked, not assembled, not pinned
// to the generic schedule, block width 2/4/8).
struct FuseKey {
  int kind, mode, lt, nb, acc;
  bool operator<(const FuseKey &o) const {
    return std::tie(kind, mode, lt, nb, acc) < std::tie(o.kind, o.mode, o.lt, o.nb, o.acc);
  }
};
static bool fuse_key(const edk_batch_item &it, int dtype, FuseKey &key) {
  edk_op *op = it.op;
  if (op->csr || op->disable_fast || !op->fast || op->fstate.load() <= 0) return false;
  if (it.nb != 2 && it.nb != 4 && it.nb != 8) return false;
  const FastState *S = (const FastState *)op->fast;
  const int var = S->tuned[dtype == EDK_C64 ? 0 : 1][it.nb];
  if (var != EDK_VARIANT_AUTO && var != EDK_VARIANT_TILE) return false;
  const FastArgs &A = S->args;
  const int mode = edk_fast_blk_mode(A);
  if (mode < 0) return false;
  const int lt = (A.L == 32 && A.A == 1) ? 32 : ((A.A == 1 && A.L >= 16 && A.kind != EDK_PROJECTED) ? 1 : 0);
  key = {A.kind, mode, lt, (int)it.nb, it.accumulate ? 1 : 0};
  return true;
}

extern "C" int edk_batch_create(edk_ctx *ctx, int dtype, const edk_batch_item *items, int64_t n, int streams,
                                edk_batch **out) {
  return edk_batch_create_ex(ctx, dtype, items, n, streams, EDK_BATCH_FUSE, out);
}

extern "C" int edk_batch_create_ex(edk_ctx *ctx, int dtype, const edk_batch_item *items, int64_t n, int streams,
                                   int flags, edk_batch **out) {
  if (!ctx || !out || n < 0 || (n > 0 && !items)) return edk_fail(EDK_ERR_ARG, "edk_batch_create: bad argument");
  if (dtype != EDK_C64 && dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "bad dtype");
  for (int64_t i = 0; i < n; i++) {
    const edk_batch_item &it = items[i];
    if (!it.op || !it.x || !it.y || it.nb < 1) return edk_fail(EDK_ERR_ARG, "batch item: bad operator/vector/nb");
    if (it.op->ctx != ctx) return edk_fail(EDK_ERR_ARG, "batch item: operator belongs to another context");
    if (it.x == it.y) return edk_fail(EDK_ERR_ARG, "batch item: x and y must not alias");
  }
  EDK_HIP(hipSetDevice(ctx->device));
  // one-time kernel setup (allocations, table builds) before capture
  for (int64_t i = 0; i < n; i++) (void)edk_matvec_prepare(items[i].op);
  EDK_HIP(hipStreamSynchronize(ctx->stream));
  if (streams <= 0) {
    // auto: a sector of >= 2^20 row-vectors fills the chip alone, and
    // running several at once only splits the last-level cache between
    // their x vectors; smaller sectors leave CUs idle, so overlap them
    bool all_large = true;
    for (int64_t i = 0; i < n; i++)
      if (items[i].op->basis->D * items[i].nb < (1 << 20)) all_large = false;
    streams = all_large ? 1 : 8;
  }
  // tasks: groups of fusable items (up to MS_MAX sectors), else single items
  std::vector<std::vector<int64_t>> tasks;
  {
    std::map<FuseKey, std::vector<int64_t>> groups;
    for (int64_t i = 0; i < n; i++) {
      FuseKey k;
      if ((flags & EDK_BATCH_FUSE) && fuse_key(items[i], dtype, k)) groups[k].push_back(i);
      else tasks.push_back({i});
    }
    for (auto &kv : groups)
      for (size_t a = 0; a < kv.second.size(); a += MS_MAX)
        tasks.emplace_back(kv.second.begin() + a, kv.second.begin() + std::min(kv.second.size(), a + MS_MAX));
  }
  auto task_load = [&](const std::vector<int64_t> &t) {
    int64_t l = 0;
    for (int64_t i : t) l += items[i].op->basis->D * items[i].nb;
    return l;
  };
  streams = std::max(1, std::min<int>(streams, (int)std::max<size_t>(tasks.size(), 1)));
  streams = std::min(streams, 16);
  // longest-processing-time assignment of tasks to branches by D * nb
  std::vector<size_t> order(tasks.size());
  for (size_t i = 0; i < tasks.size(); i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return task_load(tasks[a]) > task_load(tasks[b]); });
  std::vector<std::vector<size_t>> branch(streams);
  std::vector<int64_t> load(streams, 0);
  for (size_t t : order) {
    int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    branch[s].push_back(t);
    load[s] += task_load(tasks[t]);
  }
  edk_batch *b = new edk_batch();
  b->ctx = ctx;
  b->n = n;
  std::vector<hipStream_t> aux(streams > 1 ? streams : 0, nullptr);
  std::vector<hipEvent_t> join(aux.size(), nullptr);
  hipEvent_t fork = nullptr;
  int rc = EDK_OK;
  for (size_t s = 0; !rc && s < aux.size(); s++)
    if (hipStreamCreateWithFlags(&aux[s], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&join[s], hipEventDisableTiming) != hipSuccess)
      rc = edk_fail(EDK_ERR_HIP, "batch branch streams");
  if (!rc && !aux.empty() && hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess)
    rc = edk_fail(EDK_ERR_HIP, "batch fork event");
  bool capturing = false;
  if (!rc) {
    if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      rc = edk_fail(EDK_ERR_HIP, "hipStreamBeginCapture");
    else
      capturing = true;
  }
  if (!rc && !aux.empty()) {
    if (hipEventRecord(fork, ctx->stream) != hipSuccess) rc = edk_fail(EDK_ERR_HIP, "batch fork");
    for (size_t s = 0; !rc && s < aux.size(); s++)
      if (hipStreamWaitEvent(aux[s], fork, 0) != hipSuccess) rc = edk_fail(EDK_ERR_HIP, "batch fork wait");
  }
  for (int s = 0; !rc && s < streams; s++) {
    hipStream_t st = aux.empty() ? ctx->stream : aux[s];
    for (size_t t : branch[s]) {
      const std::vector<int64_t> &task = tasks[t];
      if (task.size() == 1) {
        const edk_batch_item &it = items[task[0]];
        if ((rc = edk_matvec_launch(it.op, dtype, it.y, it.x, it.nb, it.accumulate, st))) break;
      } else {
        // workgroup ranges padded to multiples of 8 (one per XCD)
        MultiSec M{};
        M.n = (int)task.size();
        uint32_t nbk = 0;
        for (int k = 0; k < M.n; k++) {
          const edk_batch_item &it = items[task[k]];
          M.bstart[k] = nbk;
          M.x[k] = it.x;
          M.y[k] = it.y;
          M.a[k] = ((const FastState *)it.op->fast)->args;
          nbk += (((const FastState *)it.op->fast)->args.nchunk + 7u) & ~7u;
        }
        M.bstart[M.n] = nbk;
        const edk_batch_item &i0 = items[task[0]];
        if (!edk_fast_launch_blk_multi(M, dtype, (int)i0.nb, i0.accumulate, st) ||
            hipGetLastError() != hipSuccess) {
          rc = edk_fail(EDK_ERR_HIP, "fused block launch");
          break;
        }
      }
      b->launches++;
    }
  }
  if (!rc && !aux.empty()) {
    for (size_t s = 0; !rc && s < aux.size(); s++)
      if (hipEventRecord(join[s], aux[s]) != hipSuccess || hipStreamWaitEvent(ctx->stream, join[s], 0) != hipSuccess)
        rc = edk_fail(EDK_ERR_HIP, "batch join");
  }
  if (capturing) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (!rc && e != hipSuccess) rc = edk_fail(EDK_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
  }
  if (!rc && hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
    rc = edk_fail(EDK_ERR_HIP, "hipGraphInstantiate");
  for (size_t s = 0; s < aux.size(); s++) {
    if (join[s]) (void)hipEventDestroy(join[s]);
    if (aux[s]) (void)hipStreamDestroy(aux[s]);
  }
  if (fork) (void)hipEventDestroy(fork);
  if (rc) {
    batch_free(b);
    return rc;
  }
  *out = b;
  return EDK_OK;
}

extern "C" int edk_batch_launch(edk_batch *b) {
  if (!b) return edk_fail(EDK_ERR_ARG, "null batch");
  EDK_HIP(hipSetDevice(b->ctx->device));
  if (b->n == 0) return EDK_OK;
  EDK_HIP(hipGraphLaunch(b->exec, b->ctx->stream));
  return EDK_OK;
}

extern "C" int64_t edk_batch_launches(const edk_batch *b) { return b ? b->launches : 0; }

extern "C" int edk_batch_destroy(edk_batch *b) {
  if (!b) return EDK_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  batch_free(b);
  return EDK_OK;
}
