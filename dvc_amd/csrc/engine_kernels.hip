// engine_kernels.hip -- batched frame set-up and result gathering for the
// serving engine (declarations and layouts in engine_kernels.h).
#include "engine_kernels.h"

namespace dvc {

// blockIdx.y = frame; the x blocks stride over each of its buffers.
__global__ __launch_bounds__(256) void k_job_init(const JobInitDev *items) {
  const JobInitDev it = items[blockIdx.y];
  const int t = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  for (int i = t; i < it.fs; i += stride) it.reconsTrans[i] = 0;
  for (int i = t; i < it.N; i += stride) it.carry[i] = 0xAA;
  for (int i = t; i < it.nbp; i += stride) it.state[i] = PlaneState{-1, it.numCodes - 1, 0.0f, 0, 0};
  for (int i = t; i < kMaxCodes * it.nbp; i += stride) it.rec[i] = make_int4(0, REC_NOT_RUN, -1, 0);
  if (it.accSrc)
    for (int i = t; i < it.accBytes; i += stride) it.accDst[i] = it.accSrc[i];
}

// blockIdx.y = segment.
__global__ __launch_bounds__(256) void k_copy_segments(const CopySeg *segs) {
  const CopySeg sg = segs[blockIdx.y];
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = t; i < sg.bytes; i += stride) sg.dst[i] = sg.src[i];
}

hipError_t launch_job_init(const JobInitDev *items, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_job_init, dim3(32, n), dim3(256), 0, s, items);
  return hipGetLastError();
}

hipError_t launch_copy_segments(const CopySeg *segs, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_segments, dim3(64, n), dim3(256), 0, s, segs);
  return hipGetLastError();
}

}  // namespace dvc
