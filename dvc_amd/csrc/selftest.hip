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
    case 7: ((float *)out)[i] = dvc_bp_phi_fast(((const float *)in)[i], s_logtab); break;
    default: break;
  }
}

// Every float bit pattern in [lo, hi]: the BP kernel's phi (dvc_bp_phi_fast)
// against the direct restatement of the reference's expression (dvc_bp_phi).
// counts[0] = patterns compared, counts[1] = mismatches; counts[2] = lowest
// mismatching pattern (0xffffffff when none).
__global__ void k_phi_sweep(uint32_t lo, uint32_t hi, unsigned long long *counts) {
  __shared__ double s_logtab[32];
  if (threadIdx.x < 16) {
    s_logtab[threadIdx.x] = dvcm_logf_invc[threadIdx.x];
    s_logtab[16 + threadIdx.x] = dvcm_logf_logc[threadIdx.x];
  }
  __syncthreads();
  unsigned long long n = 0, bad = 0;
  uint32_t first = 0xffffffffu;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t b = (uint64_t)lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b <= hi; b += stride) {
    const float x = __uint_as_float((uint32_t)b);
    if (x != x) continue;
    const uint32_t f = __float_as_uint(dvc_bp_phi_fast(x, s_logtab));
    const uint32_t g = __float_as_uint(dvc_bp_phi(x));
    n++;
    if (f != g) {
      bad++;
      if ((uint32_t)b < first) first = (uint32_t)b;
    }
  }
  atomicAdd(&counts[0], n);
  atomicAdd(&counts[1], bad);
  atomicMin(&counts[2], (unsigned long long)first);
}

}  // namespace dvc

extern "C" int dvc_selftest_phi_sweep(dvc_ctx *c, uint32_t lo, uint32_t hi, unsigned long long *counts) {
  DVC_CHECK(c && counts && lo <= hi, DVC_EINVAL, "dvc_selftest_phi_sweep: bad argument");
  DVC_HIP(hipSetDevice(c->device));
  unsigned long long *d = nullptr;
  const unsigned long long init[3] = {0, 0, 0xffffffffull};
  DVC_HIP(hipMalloc((void **)&d, sizeof(init)));
  hipError_t e = hipMemcpy(d, init, sizeof(init), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(dvc::k_phi_sweep, dim3(c->sm_count * 8), dim3(256), 0, c->stream, lo, hi, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(counts, d, sizeof(init), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  DVC_HIP(e);
  return DVC_OK;
}

extern "C" int dvc_selftest_unary(dvc_ctx *c, int func, const void *in, void *out, int n) {
  DVC_CHECK(c && in && out && n > 0 && func >= 0 && func <= 7, DVC_EINVAL, "dvc_selftest_unary: bad argument");
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
