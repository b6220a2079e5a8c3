// edk_runtime.hip -- contexts, errors and device-memory helpers of the C ABI.
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

#include "edk_internal.h"

static thread_local std::string g_last_error;

void edk_set_error(const std::string &msg) { g_last_error = msg; }
int edk_fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

extern "C" const char *edk_last_error(void) { return g_last_error.c_str(); }
extern "C" int edk_version(void) { return 1; }

extern "C" int edk_ctx_create(int device, void *stream, edk_ctx **out) {
  if (!out) return edk_fail(EDK_ERR_ARG, "edk_ctx_create: out is NULL");
  int n = 0;
  EDK_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return edk_fail(EDK_ERR_ARG, "edk_ctx_create: no such device");
  EDK_HIP(hipSetDevice(device));
  edk_ctx *c = new edk_ctx();
  c->device = device;
  if (stream) {
    c->stream = (hipStream_t)stream;
    c->own_stream = false;
  } else {
    // default priority on purpose: streams of different priorities are not used
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      return edk_fail(EDK_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    c->own_stream = true;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
  else c->num_cus = 256;
  hipError_t e = hipHostMalloc(&c->pinned, 2 * EDK_PINNED_SLOT * sizeof(double2));
  if (e != hipSuccess) {
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return edk_fail(EDK_ERR_OOM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  }
  // the device address of the allocation-free entry point's scratch array,
  // once per device, with this device current
  if (int rc = edk_noalloc_setup(device)) {
    (void)hipHostFree(c->pinned);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return EDK_OK;
}

// Cached allocations are rounded up to size classes (1/8 of the enclosing
// power of two above 64 MiB, at most 12.5 % extra), so a block freed by one
// job serves the next job of a similar sector.
static size_t size_class(size_t b) {
  if (b < ((size_t)64 << 20)) return (b + 255) & ~(size_t)255;
  const int e = 63 - __builtin_clzll((unsigned long long)b);
  const size_t g = (size_t)1 << (e - 3);
  return (b + g - 1) & ~(g - 1);
}

// Per-device cache of free blocks by size class.  Blocks stay cached for the
// life of the process; an allocation that finds the device full releases the
// cache and retries.
struct BlockCache {
  std::mutex mu;
  std::map<size_t, std::vector<void *>> free_blocks;
  std::unordered_map<void *, size_t> owned;  // every block the cache handed out or holds
  size_t cached = 0;
};
static std::mutex g_caches_mu;
static std::map<int, BlockCache *> g_caches;

static BlockCache &cache_of(int device) {
  std::lock_guard<std::mutex> lk(g_caches_mu);
  BlockCache *&bc = g_caches[device];
  if (!bc) bc = new BlockCache();
  return *bc;
}

static void cache_release(BlockCache &bc) {
  std::lock_guard<std::mutex> lk(bc.mu);
  for (auto &kv : bc.free_blocks)
    for (void *q : kv.second) {
      (void)hipFree(q);
      bc.owned.erase(q);
    }
  bc.free_blocks.clear();
  bc.cached = 0;
}

size_t edk_cache_bytes(int device) {
  BlockCache &bc = cache_of(device);
  std::lock_guard<std::mutex> lk(bc.mu);
  return bc.cached;
}

hipError_t edk_dalloc(edk_ctx *c, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (!c->cache_alloc) return hipMalloc(p, bytes);
  const size_t b = size_class(bytes);
  BlockCache &bc = cache_of(c->device);
  {
    std::lock_guard<std::mutex> lk(bc.mu);
    auto it = bc.free_blocks.find(b);
    if (it != bc.free_blocks.end() && !it->second.empty()) {
      *p = it->second.back();
      it->second.pop_back();
      bc.cached -= b;
      return hipSuccess;
    }
  }
  hipError_t e = hipMalloc(p, b);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    cache_release(bc);
    e = hipMalloc(p, b);
  }
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> lk(bc.mu);
    bc.owned[*p] = b;
  }
  return e;
}

void edk_dfree(edk_ctx *c, void *p) {
  if (!p) return;
  if (!c->cache_alloc) {
    (void)hipFree(p);
    return;
  }
  // the block's last use was queued on this context's stream
  (void)hipStreamSynchronize(c->stream);
  BlockCache &bc = cache_of(c->device);
  std::lock_guard<std::mutex> lk(bc.mu);
  auto it = bc.owned.find(p);
  if (it == bc.owned.end()) {
    (void)hipFree(p);
    return;
  }
  bc.free_blocks[it->second].push_back(p);
  bc.cached += it->second;
}

extern "C" int edk_ctx_destroy(edk_ctx *c) {
  if (!c) return EDK_OK;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  if (c->own_stream) hipStreamDestroy(c->stream);
  if (c->pinned) hipHostFree(c->pinned);
  delete c;
  return EDK_OK;
}

extern "C" int edk_ctx_synchronize(edk_ctx *c) {
  if (!c) return edk_fail(EDK_ERR_ARG, "null context");
  EDK_HIP(hipStreamSynchronize(c->stream));
  return EDK_OK;
}

extern "C" void *edk_ctx_stream(edk_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int edk_malloc(edk_ctx *c, size_t bytes, void **dptr) {
  if (!c || !dptr) return edk_fail(EDK_ERR_ARG, "edk_malloc: null argument");
  EDK_HIP(hipSetDevice(c->device));
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 16);
  if (e != hipSuccess) {
    *dptr = nullptr;
    return edk_fail(EDK_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  return EDK_OK;
}

extern "C" int edk_free(edk_ctx *c, void *dptr) {
  if (!c) return edk_fail(EDK_ERR_ARG, "null context");
  if (!dptr) return EDK_OK;
  EDK_HIP(hipStreamSynchronize(c->stream));
  EDK_HIP(hipFree(dptr));
  return EDK_OK;
}

extern "C" int edk_memcpy_htod(edk_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c) return edk_fail(EDK_ERR_ARG, "null context");
  if (!bytes) return EDK_OK;
  EDK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  EDK_HIP(hipStreamSynchronize(c->stream));
  return EDK_OK;
}

extern "C" int edk_memcpy_dtoh(edk_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c) return edk_fail(EDK_ERR_ARG, "null context");
  if (!bytes) return EDK_OK;
  EDK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  EDK_HIP(hipStreamSynchronize(c->stream));
  return EDK_OK;
}

extern "C" int edk_memset(edk_ctx *c, void *dptr, int value, size_t bytes) {
  if (!c) return edk_fail(EDK_ERR_ARG, "null context");
  if (!bytes) return EDK_OK;
  EDK_HIP(hipMemsetAsync(dptr, value, bytes, c->stream));
  return EDK_OK;
}
