// edk_runtime.hip -- contexts, errors and device-memory helpers of the C ABI.
#include <cstring>

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
  *out = c;
  return EDK_OK;
}

hipError_t edk_dalloc(edk_ctx *c, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (c->async_alloc) return hipMallocAsync(p, bytes, c->stream);
  return hipMalloc(p, bytes);
}

void edk_dfree(edk_ctx *c, void *p) {
  if (!p) return;
  if (c->async_alloc) (void)hipFreeAsync(p, c->stream);
  else (void)hipFree(p);
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
