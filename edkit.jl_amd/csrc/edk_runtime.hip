// edk_runtime.hip -- contexts, errors and device-memory helpers of the C ABI.
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

#include "edk_blockcache.h"
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

// The device block cache (edk_blockcache.h) over hipMalloc / hipFree, one per
// device.
struct HipRaw {
  static int alloc(void **p, size_t bytes) {
    const hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) return RAW_OK;
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? RAW_OOM : RAW_ERR;
  }
  static void free(void *p) { (void)hipFree(p); }
};
typedef BlockCacheT<HipRaw> BlockCache;
static std::mutex g_caches_mu;
static std::map<int, BlockCache *> g_caches;

static BlockCache &cache_of(int device) {
  std::lock_guard<std::mutex> lk(g_caches_mu);
  BlockCache *&bc = g_caches[device];
  if (!bc) bc = new BlockCache();
  return *bc;
}

size_t edk_cache_bytes(int device) { return (size_t)cache_of(device).stats().cached_bytes; }

hipError_t edk_dalloc(edk_ctx *c, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (!c->cache_alloc) return hipMalloc(p, bytes);
  size_t released = 0;
  const int e = cache_of(c->device).alloc(p, bytes, &released);
  if (released)
    fprintf(stderr, "[edk] device %d: block cache released (%zu bytes) after an allocation of %zu bytes failed\n",
            c->device, released, bytes);
  return e == RAW_OK ? hipSuccess : (e == RAW_OOM ? hipErrorOutOfMemory : hipErrorUnknown);
}

void edk_dfree(edk_ctx *c, void *p) {
  if (!p) return;
  if (!c->cache_alloc) {
    (void)hipFree(p);
    return;
  }
  // the block's last use was queued on this context's stream
  (void)hipStreamSynchronize(c->stream);
  cache_of(c->device).release(p);
}

extern "C" int edk_cache_info(int device, edk_cache_stats *out) {
  if (!out) return edk_fail(EDK_ERR_ARG, "edk_cache_info: null argument");
  const BlockCacheStats st = cache_of(device).stats();
  out->cached_bytes = st.cached_bytes;
  out->owned_blocks = st.owned_blocks;
  out->releases = st.releases;
  out->released_bytes = st.released_bytes;
  out->double_frees = st.double_frees;
  out->foreign_frees = st.foreign_frees;
  return EDK_OK;
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
