// noalloc.hip -- dvc_ldpca_decode_batch_noalloc: the batched speculative
// LDPCA decode with no allocation on the call.
//
// Callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc) pass device (HBM) pointers for each
// bitplane's likelihood ratios, accumulated syndrome and decoded output.
// Everything else the decode needs -- a copy of the item array, plane
// descriptors, rate-search states, run records, launch counters, LLRs and
// decision slots -- lives in `scratch`, a file-scope device array of
// DVC_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until
// the process ends.  dvc_ctx_create looks up its address on the context's
// device once (noalloc_attach); only that address is ever passed to kernels or
// copies.  The set-up and finish steps run as kernels (descriptors built and
// statistics computed on the device), so the host side allocates nothing
// either: one copy of the item array up, launches, one copy of it (with the
// statistics filled in) and an error flag back.
//
// Every call on a GPU shares that GPU's scratch: calls that overlap on one
// GPU (from two contexts, or two host threads) race.  The call synchronises
// the context's stream before it returns.
#include <mutex>
#include "dvc_internal.h"
#include "ldpca_kernels.h"

namespace dvc {

static __device__ __align__(256) uint8_t scratch[DVC_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static void *g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

// The scratch array's device address on the current device (looked up once
// per device; the device must be current).
int noalloc_attach(dvc_ctx *c) {
  DVC_CHECK(c->device >= 0 && c->device < kMaxDevices, DVC_EDEVICE, "noalloc: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[c->device]) {
    void *p = nullptr;
    DVC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DVC_CHECK(p != nullptr, DVC_EDEVICE, "noalloc: no address for the scratch array");
    g_scratch[c->device] = p;
  }
  c->noalloc_scratch = g_scratch[c->device];
  return DVC_OK;
}

static size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

// Offsets of one code-length group's buffers, after the error flag and the
// item array copy at the start of the scratch.
struct NoallocLayout {
  size_t err, items, index, planes, state, rec, counters, llr, slots, end;
};

static NoallocLayout noalloc_layout(int nTotal, int n, int N, int C) {
  NoallocLayout o;
  size_t at = 0;
  auto take = [&at](size_t b) { const size_t r = at; at = al256(at + b); return r; };
  o.err = take(sizeof(int));
  o.items = take(sizeof(dvc_ldpca_item) * (size_t)nTotal);
  o.index = take(sizeof(int) * (size_t)n);
  o.planes = take(sizeof(PlaneDesc) * (size_t)n);
  o.state = take(sizeof(PlaneState) * (size_t)n);
  o.rec = take(sizeof(int4) * (size_t)n * C);
  o.counters = take(sizeof(int) * 3);
  o.llr = take(sizeof(float) * (size_t)n * N);
  o.slots = take(((size_t)C * n + 1) * N);
  o.end = at;
  return o;
}

// One workgroup: lists the items of code length N (in item order), builds
// their plane descriptors, rate-search states and run records, and clears
// the launch counters.
__global__ __launch_bounds__(256) void noalloc_setup_kernel(const dvc_ldpca_item *items, int nTotal, int N, int C,
                                                           int *index, PlaneDesc *planes, PlaneState *state,
                                                           int4 *rec, float *llr, int *counters) {
  __shared__ int s_n;
  if (threadIdx.x == 0) {
    int k = 0;
    for (int i = 0; i < nTotal; i++)
      if (items[i].N == N) index[k++] = i;
    s_n = k;
    counters[0] = counters[1] = counters[2] = 0;
  }
  __syncthreads();
  const int n = s_n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const dvc_ldpca_item &it = items[index[j]];
    planes[j] = PlaneDesc{N, it.LR, llr + (size_t)j * N, it.acc, (uint32_t)it.crc, state + j, rec + (size_t)j * C,
                          nullptr, nullptr};
    state[j] = PlaneState{it.start_level >= 0 ? it.start_level : -1, C - 1, 0.0f, 0, 0};
  }
  for (int e = threadIdx.x; e < n * C; e += blockDim.x) rec[e] = make_int4(0, REC_NOT_RUN, -1, 0);
}

// One workgroup per bitplane of the group: its accepted decisions to the
// caller's output and its statistics (as the sequential search reports them)
// into the item array copy.  A bitplane without an accepted record in range
// sets *err.
__global__ __launch_bounds__(128) void noalloc_finish_kernel(dvc_ldpca_item *items, const int *index,
                                                            const PlaneState *state, const int4 *rec,
                                                            const uint8_t *slots, int slotCap, int N, int C,
                                                            const int *syndLength, int rateStep, int *err) {
  const int j = blockIdx.x;
  dvc_ldpca_item &it = items[index[j]];
  const PlaneState ps = state[j];
  const int4 *r = rec + (size_t)j * C;
  const int best = ps.best;
  const int4 acc = (best >= 0 && best < C) ? r[best] : make_int4(0, REC_NOT_RUN, -1, 0);
  if (acc.y != REC_ACCEPTED || acc.z < 0 || acc.z >= slotCap) {
    if (threadIdx.x == 0) atomicOr(err, 1);
    return;
  }
  const uint8_t *src = slots + (size_t)acc.z * N;
  uint8_t *dst = it.decoded;
  if (((uintptr_t)dst & 15) == 0) {
    for (int i = threadIdx.x; i < N / 16; i += blockDim.x) ((uint4 *)dst)[i] = ((const uint4 *)src)[i];
  } else {
    for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = src[i];
  }
  if (threadIdx.x != 0) return;
  dvc_bp_stats st;
  st.start_level = ps.start;
  st.entropy = ps.entropy;
  int lvl = -1, iters = 0;
  for (int l = ps.start; l < C; l++) {
    if (r[l].y == REC_ACCEPTED) { lvl = l; break; }
    iters += r[l].x;
  }
  if (lvl < 0) { atomicOr(err, 1); return; }
  st.level = lvl;
  st.iterations = iters + r[lvl].x;
  st.rate = syndLength[lvl] + 8;
  if (rateStep) {
    st.requests = 1;
    for (int off = 0; ps.start + off < lvl; off = dynstep_next(off)) st.requests++;
  } else {
    st.requests = lvl - ps.start + 1;
  }
  st.used_inverse = r[lvl].w;
  st.iterations_run = 0;
  for (int l = 0; l < C; l++)
    if (r[l].y != REC_NOT_RUN) st.iterations_run += r[l].x;
  it.stats = st;
}

}  // namespace dvc

using namespace dvc;

extern "C" int dvc_ldpca_decode_batch_noalloc(dvc_ctx *c, int n, dvc_ldpca_item *items) {
  DVC_CHECK(c, DVC_EINVAL, "dvc_ldpca_decode_batch_noalloc: NULL context");
  DVC_CHECK(n >= 0 && (n == 0 || items), DVC_EINVAL, "dvc_ldpca_decode_batch_noalloc: bad item list");
  DVC_CHECK(c->bp_mode == 0, DVC_EINVAL,
            "dvc_ldpca_decode_batch_noalloc: offered with parallel_mode 0 only (ParallelMode 1 carries decisions "
            "between codes through the host-driven resolution)");
  DVC_CHECK(c->noalloc_scratch, DVC_ESTATE, "dvc_ldpca_decode_batch_noalloc: scratch not attached");
  if (n == 0) return DVC_OK;
  int count[2] = {0, 0};
  for (int i = 0; i < n; i++) {
    DVC_CHECK(items[i].N == 1584 || items[i].N == 6336, DVC_EINVAL,
              "dvc_ldpca_decode_batch_noalloc: N must be 1584 or 6336");
    DVC_CHECK(items[i].LR && items[i].acc && items[i].decoded, DVC_EINVAL,
              "dvc_ldpca_decode_batch_noalloc: NULL item pointer");
    DVC_CHECK(items[i].start_level < c->ladder(items[i].N)->numCodes, DVC_EINVAL,
              "dvc_ldpca_decode_batch_noalloc: start_level beyond the ladder");
    count[items[i].N == 6336]++;
  }
  for (int g = 0; g < 2; g++) {
    const int N = g ? 6336 : 1584;
    const NoallocLayout o = noalloc_layout(n, count[g], N, c->ladder(N)->numCodes);
    DVC_CHECK(o.end <= (size_t)DVC_NOALLOC_SCRATCH_BYTES, DVC_EINVAL,
              "dvc_ldpca_decode_batch_noalloc: the bitplanes need more than the fixed scratch of "
              "DVC_NOALLOC_SCRATCH_BYTES (at most 150 bitplanes of N = 6336 or 600 of N = 1584 per call)");
  }
  DVC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  uint8_t *base = (uint8_t *)c->noalloc_scratch;
  const NoallocLayout o0 = noalloc_layout(n, 0, 1584, 65);
  int *dErr = (int *)(base + o0.err);
  dvc_ldpca_item *dItems = (dvc_ldpca_item *)(base + o0.items);
  DVC_HIP(launch_zero_ints(dErr, 1, s));
  DVC_HIP(hipMemcpyAsync(dItems, items, sizeof(dvc_ldpca_item) * (size_t)n, hipMemcpyHostToDevice, s));
  for (int g = 0; g < 2; g++) {
    const int ng = count[g];
    if (!ng) continue;
    const int N = g ? 6336 : 1584;
    Ladder *L = c->ladder(N);
    const int C = L->numCodes, maxOff = C - 1, slotCap = C * ng + 1;
    const NoallocLayout o = noalloc_layout(n, ng, N, C);
    int *dIndex = (int *)(base + o.index), *dCounters = (int *)(base + o.counters);
    PlaneDesc *dPlanes = (PlaneDesc *)(base + o.planes);
    PlaneState *dState = (PlaneState *)(base + o.state);
    int4 *dRec = (int4 *)(base + o.rec);
    float *dLLR = (float *)(base + o.llr);
    uint8_t *dSlots = base + o.slots;
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3(1), dim3(256), 0, s, dItems, n, N, C, dIndex, dPlanes, dState, dRec,
                       dLLR, dCounters);
    DVC_HIP(hipGetLastError());
    DVC_HIP(launch_ce_llr(dPlanes, ng, 1, c->lad1584, c->lad6336, s));
    const LadderDev ld = ladder_dev(*L);
    BpArgs a{ld, dPlanes, ng, dCounters, ng * maxOff, maxOff, c->max_iter, dSlots, dCounters + 1, slotCap,
             c->bp_poll, c->bp_mode, c->rate_step};
    DVC_HIP(launch_bp_any(N, bp_threads(c, N), a, bp_grid(c, N, ng * maxOff), s));
    DVC_HIP(launch_invh(dPlanes, ng, ld, dSlots, dCounters + 1, slotCap, s));
    hipLaunchKernelGGL(noalloc_finish_kernel, dim3(ng), dim3(128), 0, s, dItems, dIndex, dState, dRec, dSlots, slotCap,
                       N, C, L->d_syndLength, c->rate_step, dErr);
    DVC_HIP(hipGetLastError());
  }
  int err = 0;
  DVC_HIP(hipMemcpyAsync(items, dItems, sizeof(dvc_ldpca_item) * (size_t)n, hipMemcpyDeviceToHost, s));
  DVC_HIP(hipMemcpyAsync(&err, dErr, sizeof(int), hipMemcpyDeviceToHost, s));
  DVC_HIP(hipStreamSynchronize(s));
  DVC_CHECK(err == 0, DVC_ESTATE, "dvc_ldpca_decode_batch_noalloc: a bitplane has no accepted code in range");
  return DVC_OK;
}
