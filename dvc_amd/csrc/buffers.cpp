// buffers.cpp -- device buffers for callers without a HIP runtime of their
// own (the Python binding, tests, benchmarks): allocate, free and copy on the
// context's device and stream.
#include "dvc_internal.h"

using namespace dvc;

extern "C" {

int dvc_buffer_alloc(dvc_ctx *c, size_t bytes, void **dptr) {
  DVC_CHECK(c && dptr, DVC_EINVAL, "dvc_buffer_alloc: NULL argument");
  DVC_HIP(hipSetDevice(c->device));
  *dptr = nullptr;
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    set_error(std::string("dvc_buffer_alloc: ") + hipGetErrorString(e));
    return DVC_ENOMEM;
  }
  return DVC_OK;
}

int dvc_buffer_free(dvc_ctx *c, void *dptr) {
  DVC_CHECK(c, DVC_EINVAL, "dvc_buffer_free: NULL context");
  if (!dptr) return DVC_OK;
  DVC_HIP(hipSetDevice(c->device));
  DVC_HIP(hipFree(dptr));
  return DVC_OK;
}

// kind: 0 host -> device, 1 device -> host, 2 device -> device.  Returns once
// the copy is complete.
int dvc_buffer_copy(dvc_ctx *c, void *dst, const void *src, size_t bytes, int kind) {
  DVC_CHECK(c && dst && src && kind >= 0 && kind <= 2, DVC_EINVAL, "dvc_buffer_copy: bad argument");
  DVC_HIP(hipSetDevice(c->device));
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  DVC_HIP(hipMemcpyAsync(dst, src, bytes, k, c->stream));
  DVC_HIP(hipStreamSynchronize(c->stream));
  return DVC_OK;
}

}  // extern "C"
