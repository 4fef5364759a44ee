// selftest.hip -- device evaluation of the scalar routines the kernels rely
// on, so tests can compare them with the host C library on the GPU itself.
#include "dvc_internal.h"
#include "dvc_math.h"

#pragma clang fp contract(off)

namespace dvc {

__global__ void k_selftest(int func, const void *in, void *out, int n) {
  __shared__ double s_logtab[32];  // as the BP kernel stages it
  if (threadIdx.x < 16) {
    s_logtab[threadIdx.x] = dvcm_logf_invc[threadIdx.x];
    s_logtab[16 + threadIdx.x] = dvcm_logf_logc[threadIdx.x];
  }
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  switch (func) {
    case 0: ((float *)out)[i] = dvc_logf(((const float *)in)[i]); break;
    case 1: ((float *)out)[i] = dvc_expf(((const float *)in)[i]); break;
    case 2: ((float *)out)[i] = dvc_tanhf(((const float *)in)[i]); break;
    case 3: ((float *)out)[i] = dvc_bp_phi(((const float *)in)[i]); break;
    case 4: ((double *)out)[i] = sqrt((double)((const int *)in)[i]); break;
    case 5: ((double *)out)[i] = exp(((const double *)in)[i]); break;
    case 6: ((float *)out)[i] = dvc_bp_phi_nonneg(((const float *)in)[i], s_logtab); break;
    default: break;
  }
}

}  // namespace dvc

extern "C" int dvc_selftest_unary(dvc_ctx *c, int func, const void *in, void *out, int n) {
  DVC_CHECK(c && in && out && n > 0 && func >= 0 && func <= 6, DVC_EINVAL, "dvc_selftest_unary: bad argument");
  DVC_HIP(hipSetDevice(c->device));
  const size_t isz = func == 5 ? 8 : 4, osz = (func == 4 || func == 5) ? 8 : 4;
  void *din = nullptr, *dout = nullptr;
  DVC_HIP(hipMalloc(&din, isz * (size_t)n));
  hipError_t e = hipMalloc(&dout, osz * (size_t)n);
  if (e != hipSuccess) { (void)hipFree(din); dvc::set_error("dvc_selftest_unary: hipMalloc"); return DVC_ENOMEM; }
  e = hipMemcpy(din, in, isz * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(dvc::k_selftest, dim3((n + 255) / 256), dim3(256), 0, c->stream, func, din, dout, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(out, dout, osz * (size_t)n, hipMemcpyDeviceToHost);
  (void)hipFree(din);
  (void)hipFree(dout);
  DVC_HIP(e);
  return DVC_OK;
}
