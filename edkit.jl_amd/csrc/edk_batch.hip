// edk_batch.hip -- multi-sector matvec batches as one HIP graph, and the
// schedule autotuner's C ABI.
//
// A batch records the matvecs of its items (operator, x, y, block width) once
// under stream capture and replays them with a single hipGraphLaunch.  Items
// are spread over `streams` capture branches (fork from the context stream,
// join back), balanced by D * nb, so the graph has that many independent
// chains and small sectors do not wait for each other.  Capture records no
// memset and no allocation: every kernel writes all of its y (accumulate = 0)
// or reads it (accumulate = 1), and the kernels' one-time setup runs before
// capture begins.
#include <algorithm>
#include <vector>

#include "edk_matvec.h"

struct edk_batch {
  edk_ctx *ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t n = 0;
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

extern "C" int edk_batch_create(edk_ctx *ctx, int dtype, const edk_batch_item *items, int64_t n, int streams,
                                edk_batch **out) {
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
  streams = std::max(1, std::min<int>(streams, (int)std::max<int64_t>(n, 1)));
  streams = std::min(streams, 16);
  // longest-processing-time assignment of items to branches by D * nb
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    return items[a].op->basis->D * items[a].nb > items[b].op->basis->D * items[b].nb;
  });
  std::vector<std::vector<int64_t>> branch(streams);
  std::vector<int64_t> load(streams, 0);
  for (int64_t i : order) {
    int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    branch[s].push_back(i);
    load[s] += items[i].op->basis->D * items[i].nb;
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
    for (int64_t i : branch[s]) {
      const edk_batch_item &it = items[i];
      if ((rc = edk_matvec_launch(it.op, dtype, it.y, it.x, it.nb, it.accumulate, st))) break;
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

extern "C" int edk_batch_destroy(edk_batch *b) {
  if (!b) return EDK_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  batch_free(b);
  return EDK_OK;
}
