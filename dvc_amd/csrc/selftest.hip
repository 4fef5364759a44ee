// selftest.hip -- device evaluation of the scalar routines the kernels rely
// on, so tests can compare them with the host C library on the GPU itself.
#include "dvc_internal.h"
#include "dvc_math.h"
#include <vector>

#pragma clang fp contract(off)

namespace dvc {

__global__ void k_selftest(int func, const void *in, void *out, int n) {
  __shared__ double s_logtab[32];  // dvc_logf_normal's (invc, logc) table
  __shared__ __align__(16) double s_phitab[2 * DVC_PHI_LOGTAB];  // as the BP kernels stage it
  if (threadIdx.x < 16) {
    s_logtab[threadIdx.x] = dvcm_logf_invc[threadIdx.x];
    s_logtab[16 + threadIdx.x] = dvcm_logf_logc[threadIdx.x];
  }
  for (int e = threadIdx.x; e < DVC_PHI_LOGTAB; e += blockDim.x) dvc_phi_logtab_entry(e, s_phitab);
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  switch (func) {
    case 0: ((float *)out)[i] = dvc_logf(((const float *)in)[i]); break;
    case 1: ((float *)out)[i] = dvc_expf(((const float *)in)[i]); break;
    case 2: ((float *)out)[i] = dvc_tanhf(((const float *)in)[i]); break;
    case 3: ((float *)out)[i] = dvc_bp_phi(((const float *)in)[i]); break;
    case 4: ((double *)out)[i] = sqrt((double)((const int *)in)[i]); break;
    case 5: ((double *)out)[i] = dvc_exp(((const double *)in)[i]); break;
    case 6: ((float *)out)[i] = dvc_bp_phi_nonneg(((const float *)in)[i], s_logtab); break;
    case 7: ((float *)out)[i] = dvc_bp_phi_fast(((const float *)in)[i], s_phitab); break;
    default: break;
  }
}

// Every float bit pattern in [lo, hi]: the BP kernel's phi (dvc_bp_phi_fast)
// against the direct restatement of the reference's expression (dvc_bp_phi).
// counts[0] = patterns compared, counts[1] = mismatches; counts[2] = lowest
// mismatching pattern (0xffffffff when none).
__global__ void k_phi_sweep(uint32_t lo, uint32_t hi, unsigned long long *counts) {
  __shared__ __align__(16) double s_phitab[2 * DVC_PHI_LOGTAB];
  for (int e = threadIdx.x; e < DVC_PHI_LOGTAB; e += blockDim.x) dvc_phi_logtab_entry(e, s_phitab);
  __syncthreads();
  unsigned long long n = 0, bad = 0;
  uint32_t first = 0xffffffffu;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t b = (uint64_t)lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b <= hi; b += stride) {
    const float x = __uint_as_float((uint32_t)b);
    if (x != x) continue;
    const uint32_t f = __float_as_uint(dvc_bp_phi_fast(x, s_phitab));
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

// VALU issue-rate calibration: 8 independent chains per lane, 128 operations
// of one kind per trip (kind 0: v_fma_f32, 1: v_add_f32, 2: v_mul_f32, 3:
// v_fma_f64, 4: v_add_u32, 5: v_sad_u8 (4 absolute differences), 6:
// v_mqsad_u32_u8 (16), 7: v_qsad_pk_u16_u8 (16), 8: v_rcp_f32, a
// transcendental), built with the library's flags (packed FP32 off).
// The loop's own counter is scalar, so a trip issues 128 VALU instructions
// per wave.
constexpr int VALU_CHAINS = 8, VALU_UNROLL = 16;
template <int KIND>
__global__ __launch_bounds__(256) void k_valu_rate(const float *in, float *out, int iters, float a, float b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if constexpr (KIND == 3) {
    double x[VALU_CHAINS];
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) x[q] = (double)in[i & 1023] + q;
    for (int it = 0; it < iters; it++) {
#pragma unroll
      for (int u = 0; u < VALU_UNROLL; u++)
#pragma unroll
        for (int q = 0; q < VALU_CHAINS; q++) x[q] = __builtin_fma(x[q], (double)a, (double)b);
    }
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) t += x[q];
    out[i] = (float)t;
  } else if constexpr (KIND == 5) {
    uint32_t x[VALU_CHAINS];
    const uint32_t k = __float_as_uint(a);
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) x[q] = __float_as_uint(in[i & 1023]) + q;
    for (int it = 0; it < iters; it++) {
#pragma unroll
      for (int u = 0; u < VALU_UNROLL; u++)
#pragma unroll
        for (int q = 0; q < VALU_CHAINS; q++) x[q] = __builtin_amdgcn_sad_u8(x[q], k, x[q]);
    }
    uint32_t t = 0;
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) t ^= x[q];
    out[i] = __uint_as_float(t);
  } else if constexpr (KIND == 6) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    u32x4 acc[VALU_CHAINS];
    const uint64_t src = ((uint64_t)__float_as_uint(b) << 32) | __float_as_uint(in[i & 1023]);
    const uint32_t k = __float_as_uint(a);
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) acc[q] = u32x4{(uint32_t)q, 0u, 0u, 0u};
    for (int it = 0; it < iters; it++) {
#pragma unroll
      for (int u = 0; u < VALU_UNROLL; u++)
#pragma unroll
        for (int q = 0; q < VALU_CHAINS; q++) acc[q] = __builtin_amdgcn_mqsad_u32_u8(src, k, acc[q]);
    }
    uint32_t t = 0;
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) t ^= acc[q].x ^ acc[q].y ^ acc[q].z ^ acc[q].w;
    out[i] = __uint_as_float(t);
  } else if constexpr (KIND == 7) {
    uint64_t acc[VALU_CHAINS];
    const uint64_t src = ((uint64_t)__float_as_uint(b) << 32) | __float_as_uint(in[i & 1023]);
    const uint32_t k = __float_as_uint(a);
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) acc[q] = (uint64_t)q;
    for (int it = 0; it < iters; it++) {
#pragma unroll
      for (int u = 0; u < VALU_UNROLL; u++)
#pragma unroll
        for (int q = 0; q < VALU_CHAINS; q++) acc[q] = __builtin_amdgcn_qsad_pk_u16_u8(src, k, acc[q]);
    }
    uint64_t t = 0;
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) t ^= acc[q];
    out[i] = __uint_as_float((uint32_t)t ^ (uint32_t)(t >> 32));
  } else if constexpr (KIND == 8) {
    float x[VALU_CHAINS];
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) x[q] = in[i & 1023] + (float)q;
    for (int it = 0; it < iters; it++) {
#pragma unroll
      for (int u = 0; u < VALU_UNROLL; u++)
#pragma unroll
        for (int q = 0; q < VALU_CHAINS; q++) x[q] = __builtin_amdgcn_rcpf(x[q]);
    }
    float t = 0.0f;
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) t += x[q];
    out[i] = t;
  } else if constexpr (KIND == 4) {
    uint32_t x[VALU_CHAINS];
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) x[q] = __float_as_uint(in[i & 1023]) + q + __float_as_uint(a);
    for (int it = 0; it < iters; it++) {
#pragma unroll
      for (int u = 0; u < VALU_UNROLL; u++)
#pragma unroll
        for (int q = 0; q < VALU_CHAINS; q++) x[q] = x[q] + x[(q + 1) % VALU_CHAINS];
    }
    uint32_t t = 0;
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) t ^= x[q];
    out[i] = __uint_as_float(t);
  } else {
    float x[VALU_CHAINS];
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) x[q] = in[i & 1023] + (float)q;
    for (int it = 0; it < iters; it++) {
#pragma unroll
      for (int u = 0; u < VALU_UNROLL; u++)
#pragma unroll
        for (int q = 0; q < VALU_CHAINS; q++) {
          if constexpr (KIND == 0) x[q] = __builtin_fmaf(x[q], a, b);
          else if constexpr (KIND == 1) x[q] = x[q] + a;
          else x[q] = x[q] * a;
        }
    }
    float t = 0.0f;
#pragma unroll
    for (int q = 0; q < VALU_CHAINS; q++) t += x[q];
    out[i] = t;
  }
}

}  // namespace dvc

// Verification entry (no counterpart in the reference): times one launch of
// the VALU calibration kernel of `kind` (0 fp32 fma, 1 fp32 add, 2 fp32 mul,
// 3 fp64 fma, 4 int32 add, 5 sad_u8, 6 mqsad_u32_u8, 7 qsad_pk_u16_u8, 8 rcp_f32)
// over `waves_per_cu` waves per CU and `iters`
// trips; writes the kernel time (HIP events on the context stream) and the
// VALU wave-instructions it issues (128 per trip per wave).
extern "C" int dvc_selftest_valu_rate(dvc_ctx *c, int kind, int waves_per_cu, int iters, double *ms,
                                      double *wave_instr) {
  DVC_CHECK(c && ms && wave_instr && kind >= 0 && kind <= 8 && waves_per_cu >= 4 && waves_per_cu % 4 == 0 &&
                waves_per_cu <= 32 && iters > 0,
            DVC_EINVAL, "dvc_selftest_valu_rate: bad argument");
  DVC_HIP(hipSetDevice(c->device));
  const int blocks = c->sm_count * waves_per_cu / 4, threads = 256, n = blocks * threads;
  float *din = nullptr, *dout = nullptr;
  DVC_HIP(hipMalloc((void **)&din, 1024 * sizeof(float)));
  hipError_t e = hipMalloc((void **)&dout, sizeof(float) * (size_t)n);
  std::vector<float> init(1024);
  for (int k = 0; k < 1024; k++) init[k] = 1.0f + 1e-3f * (float)k;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipMemcpy(din, init.data(), init.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (int rep = 0; rep < 2 && e == hipSuccess; rep++) {  // first launch warms up, second is timed
    if (rep == 1) e = hipEventRecord(e0, c->stream);
    const float a = 0.999f, b = 1e-3f;
    switch (kind) {
      case 0: hipLaunchKernelGGL(dvc::k_valu_rate<0>, dim3(blocks), dim3(threads), 0, c->stream, din, dout, iters, a, b); break;
      case 1: hipLaunchKernelGGL(dvc::k_valu_rate<1>, dim3(blocks), dim3(threads), 0, c->stream, din, dout, iters, a, b); break;
      case 2: hipLaunchKernelGGL(dvc::k_valu_rate<2>, dim3(blocks), dim3(threads), 0, c->stream, din, dout, iters, a, b); break;
      case 3: hipLaunchKernelGGL(dvc::k_valu_rate<3>, dim3(blocks), dim3(threads), 0, c->stream, din, dout, iters, a, b); break;
      case 4: hipLaunchKernelGGL(dvc::k_valu_rate<4>, dim3(blocks), dim3(threads), 0, c->stream, din, dout, iters, a, b); break;
      case 5: hipLaunchKernelGGL(dvc::k_valu_rate<5>, dim3(blocks), dim3(threads), 0, c->stream, din, dout, iters, a, b); break;
      case 6: hipLaunchKernelGGL(dvc::k_valu_rate<6>, dim3(blocks), dim3(threads), 0, c->stream, din, dout, iters, a, b); break;
      case 7: hipLaunchKernelGGL(dvc::k_valu_rate<7>, dim3(blocks), dim3(threads), 0, c->stream, din, dout, iters, a, b); break;
      default: hipLaunchKernelGGL(dvc::k_valu_rate<8>, dim3(blocks), dim3(threads), 0, c->stream, din, dout, iters, a, b); break;
    }
    e = hipGetLastError();
    if (e == hipSuccess && rep == 1) e = hipEventRecord(e1, c->stream);
  }
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(din);
  if (dout) (void)hipFree(dout);
  DVC_HIP(e);
  *ms = t;
  *wave_instr = (double)blocks * (threads / 64) * (double)iters * dvc::VALU_CHAINS * dvc::VALU_UNROLL;
  return DVC_OK;
}

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
