// capi_ldpca.cpp -- context management and the LDPCA entry points of
// include/dvc_accel.h.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <dlfcn.h>
#include <string>
#include <vector>
#include <algorithm>
#include "dvc_internal.h"
#include "ldpca_kernels.h"

namespace dvc {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

// Directory of this shared library's data files: <lib dir>/data.
static std::string default_data_dir() {
  Dl_info info;
  if (dladdr((void *)&default_data_dir, &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    size_t s = p.find_last_of('/');
    if (s != std::string::npos) return p.substr(0, s) + "/data";
  }
  return "data";
}

int bp_threads(const dvc_ctx *c, int N) { return N == 1584 ? c->bp_threads_qcif : c->bp_threads_cif; }

int bp_grid(const dvc_ctx *c, int N, int items) {
  // a CIF codeword takes 118 KB of a CU's 160 KB LDS; QCIF codewords (30 KB) share one
  const int perCU = N == 6336 ? 1 : c->bp_wg_per_cu_qcif;
  return std::max(1, std::min(c->sm_count * perCU, items));
}

}  // namespace dvc

using namespace dvc;

hipEvent_t dvc_ctx::timer_begin(const char *name, hipStream_t s, int *slot) {
  *slot = -1;
  if (!profiling) return nullptr;
  std::lock_guard<std::mutex> lock(timer_mu);
  int k = -1;
  for (size_t i = 0; i < timers.size(); i++)
    if (timers[i].name == name) { k = (int)i; break; }
  if (k < 0) { timers.push_back(KernelTimer{name}); k = (int)timers.size() - 1; }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  (void)hipEventRecord(e, s);
  *slot = k;
  return e;
}

void dvc_ctx::timer_end(int slot, hipEvent_t start, hipStream_t s) {
  if (slot < 0 || !start) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  std::lock_guard<std::mutex> lock(timer_mu);
  timers[slot].pending.push_back({start, e});
}

void dvc_ctx::timer_collect() {
  std::lock_guard<std::mutex> lock(timer_mu);
  for (auto &t : timers) {
    for (auto &pe : t.pending) {
      float ms = 0.0f;
      if (hipEventSynchronize(pe.second) == hipSuccess && hipEventElapsedTime(&ms, pe.first, pe.second) == hipSuccess) {
        t.total_ms += ms;
        t.launches++;
      }
      (void)hipEventDestroy(pe.first);
      (void)hipEventDestroy(pe.second);
    }
    t.pending.clear();
  }
}

// Timed launch helper: TIMED(ctx, "name", stream, expr-returning-hipError_t)
#define TIMED(c, name, s, expr)                                  \
  do {                                                           \
    int slot_;                                                   \
    hipEvent_t ev_ = (c)->timer_begin(name, s, &slot_);          \
    DVC_HIP(expr);                                               \
    (c)->timer_end(slot_, ev_, s);                               \
  } while (0)

extern "C" {

const char *dvc_last_error(void) { return g_err.c_str(); }

int dvc_ctx_create(const dvc_params *params, dvc_ctx **out) {
  DVC_CHECK(out != nullptr, DVC_EINVAL, "dvc_ctx_create: out is NULL");
  *out = nullptr;
  dvc_params p{};
  if (params) p = *params;
  DVC_CHECK(p.parallel_mode == 0 || p.parallel_mode == 1, DVC_EINVAL, "dvc_ctx_create: parallel_mode must be 0 or 1");
  DVC_CHECK(p.rate_step == 0 || (p.rate_step == 1 && p.parallel_mode == 0), DVC_EINVAL,
            "dvc_ctx_create: rate_step must be 0, or 1 with parallel_mode 0");
  DVC_CHECK(p.qcif_ladder == 0 || p.qcif_ladder == 1, DVC_EINVAL, "dvc_ctx_create: qcif_ladder must be 0 or 1");
  int ndev = 0;
  DVC_HIP(hipGetDeviceCount(&ndev));
  DVC_CHECK(ndev > 0 && p.device >= 0 && p.device < ndev, DVC_EDEVICE, "dvc_ctx_create: no such HIP device");
  DVC_HIP(hipSetDevice(p.device));
  hipDeviceProp_t prop;
  DVC_HIP(hipGetDeviceProperties(&prop, p.device));
  DVC_CHECK(strncmp(prop.gcnArchName, "gfx950", 6) == 0, DVC_EDEVICE,
            std::string("dvc_ctx_create: this library is built for gfx950, device is ") + prop.gcnArchName);
  dvc_ctx *c = new dvc_ctx();
  c->device = p.device;
  c->max_iter = p.max_iter > 0 ? p.max_iter : 50;
  if (p.bp_threads > 0) c->bp_threads_cif = p.bp_threads;
  c->bp_mode = p.parallel_mode;
  c->rate_step = p.rate_step;
  c->sm_count = prop.multiProcessorCount;
  std::string dir = p.data_dir ? std::string(p.data_dir) : default_data_dir();
  int rc;
  if ((rc = load_ladder(dir, p.qcif_ladder ? "irreg" : "reg", 1584, c->lad1584)) ||
      (rc = load_ladder(dir, "reg", 6336, c->lad6336)) ||
      (rc = upload_ladder(c->lad1584)) || (rc = upload_ladder(c->lad6336))) {
    dvc_ctx_destroy(c);
    return rc;
  }
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { dvc_ctx_destroy(c); set_error("hipStreamCreate failed"); return DVC_EDEVICE; }
  if ((rc = noalloc_attach(c))) {
    dvc_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return DVC_OK;
}

int dvc_ctx_destroy(dvc_ctx *c) {
  if (!c) return DVC_OK;
  (void)hipSetDevice(c->device);
  free_ladder(c->lad1584);
  free_ladder(c->lad6336);
  if (c->d_costFme) (void)hipFree(c->d_costFme);
  if (c->d_costUsi) (void)hipFree(c->d_costUsi);
  if (c->si_scratch) (void)hipFree(c->si_scratch);
  if (c->ld_dev) (void)hipFree(c->ld_dev);
  if (c->ld_host) (void)hipHostFree(c->ld_host);
  band_free(c);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return DVC_OK;
}

int dvc_prepare_data(const char *data_dir) {
  std::string dir = data_dir ? std::string(data_dir) : default_data_dir();
  const std::pair<const char *, int> packs[] = {{"reg", 1584}, {"reg", 6336}, {"irreg", 1584}};
  for (const auto &pk : packs) {
    Ladder L;
    int rc = load_ladder(dir, pk.first, pk.second, L);
    if (rc) return rc;
  }
  return DVC_OK;
}

int dvc_profile(dvc_ctx *c, int enable) {
  DVC_CHECK(c, DVC_EINVAL, "dvc_profile: NULL context");
  c->timer_collect();
  c->timers.clear();
  c->profiling = enable != 0;
  return DVC_OK;
}

int dvc_profile_get(dvc_ctx *c, const char *name, double *total_ms, long *launches) {
  DVC_CHECK(c && name, DVC_EINVAL, "dvc_profile_get: NULL argument");
  c->timer_collect();
  *total_ms = 0.0;
  *launches = 0;
  for (auto &t : c->timers)
    if (t.name == name) { *total_ms = t.total_ms; *launches = t.launches; }
  return DVC_OK;
}

int dvc_ladder_info(dvc_ctx *c, int N, int *num_codes, int *num_edges, int *synd_length) {
  DVC_CHECK(c, DVC_EINVAL, "dvc_ladder_info: NULL context");
  Ladder *L = c->ladder(N);
  DVC_CHECK(L, DVC_EINVAL, "dvc_ladder_info: N must be 1584 or 6336");
  if (num_codes) *num_codes = L->numCodes;
  if (num_edges) *num_edges = L->E;
  if (synd_length) memcpy(synd_length, L->syndLength.data(), sizeof(int) * L->numCodes);
  return DVC_OK;
}

int dvc_ladder_export(dvc_ctx *c, int N, int code, int *ir_out, uint8_t *inv_out) {
  DVC_CHECK(c, DVC_EINVAL, "dvc_ladder_export: NULL context");
  Ladder *L = c->ladder(N);
  DVC_CHECK(L && code >= 0 && code < L->numCodes, DVC_EINVAL, "dvc_ladder_export: bad N or code");
  if (ir_out) memcpy(ir_out, &L->decIr[(size_t)code * L->E], sizeof(int) * L->E);
  if (inv_out) memcpy(inv_out, L->invH.data(), L->invH.size());
  return DVC_OK;
}

// Byte offsets of decode_group's buffers in the context's LDPCA workspace.
// The input region [lr, inEnd) is filled in pinned host memory at the same
// offsets and uploaded in one copy; [state, outEnd) comes back in one copy.
struct LdpcaLayout {
  size_t lr, acc, carry, planes, state, rec, counters, inEnd, out, outEnd, llr, slots, bcase, total;
};

static LdpcaLayout ldpca_layout(size_t n, size_t N, size_t C, size_t slotCap, bool mode1) {
  LdpcaLayout o;
  size_t at = 0;
  auto take = [&at](size_t bytes) {
    const size_t r = at;
    at += (bytes + 255) & ~(size_t)255;
    return r;
  };
  o.lr = take(sizeof(float) * n * N);
  o.acc = take(n * (N / 8));
  o.carry = take(mode1 ? n * N : 0);
  o.planes = take(sizeof(PlaneDesc) * n);
  o.state = take(sizeof(PlaneState) * n);
  o.rec = take(sizeof(int4) * n * C);
  o.counters = take(sizeof(int) * 3);
  o.inEnd = at;
  o.out = take(n * N);
  o.outEnd = o.out + n * N;
  o.llr = take(sizeof(float) * n * N);
  o.slots = take(slotCap * N);
  o.bcase = take(mode1 ? sizeof(int) * n * C : 0);
  o.total = at;
  return o;
}

// Grows the context's LDPCA workspace (device) and its pinned staging (host);
// both are kept between calls and freed by dvc_ctx_destroy.
static int ldpca_workspace(dvc_ctx *c, size_t devBytes, size_t hostBytes) {
  if (c->ld_dev_bytes < devBytes) {
    if (c->ld_dev) DVC_HIP(hipFree(c->ld_dev));
    c->ld_dev = nullptr;
    c->ld_dev_bytes = 0;
    DVC_HIP(hipMalloc(&c->ld_dev, devBytes));
    c->ld_dev_bytes = devBytes;
  }
  if (c->ld_host_bytes < hostBytes) {
    if (c->ld_host) DVC_HIP(hipHostFree(c->ld_host));
    c->ld_host = nullptr;
    c->ld_host_bytes = 0;
    DVC_HIP(hipHostMalloc(&c->ld_host, hostBytes, hipHostMallocDefault));
    c->ld_host_bytes = hostBytes;
  }
  return DVC_OK;
}

// Shared driver for one group of bitplanes with the same code length.
// sequential != 0: one code per launch in increasing order (the plain path,
// which mirrors the reference's per-code beliefPropagation_CUDA call).
// ParallelMode 1: items[i].decoded holds the decisions carried into the
// bitplane's first code on entry (the reference's in/out decoded buffer).
static int decode_group(dvc_ctx *c, int N, int n, dvc_ldpca_item *items, int sequential, hipStream_t s) {
  Ladder *L = c->ladder(N);
  const int C = L->numCodes, nb = N / 8;
  const bool mode1 = c->bp_mode == 1;
  // slots: each (bitplane, code) accepts at most once; ParallelMode 1 keeps
  // every run code's decisions and its repeat-case decisions as well
  const int slotCap = sequential ? (mode1 ? 2 * C + 2 : 2) : (mode1 ? 2 * C * n + 1 : C * n + 1);
  const LdpcaLayout o = ldpca_layout(n, N, C, slotCap, mode1);
  int rc = ldpca_workspace(c, o.total, o.outEnd);
  if (rc) return rc;
  uint8_t *dev = (uint8_t *)c->ld_dev, *host = (uint8_t *)c->ld_host;
  float *dLR = (float *)(dev + o.lr), *dLLR = (float *)(dev + o.llr);
  uint8_t *dAcc = dev + o.acc, *dCarry = dev + o.carry, *dSlots = dev + o.slots, *dOut = dev + o.out;
  PlaneDesc *dPlanes = (PlaneDesc *)(dev + o.planes);
  PlaneState *dState = (PlaneState *)(dev + o.state);
  int4 *dRec = (int4 *)(dev + o.rec);
  int *dCounters = (int *)(dev + o.counters);  // [0] queue head, [1] slot count, [2] bitplanes to resume
  int *dBcase = (int *)(dev + o.bcase);
  PlaneDesc *hPlanes = (PlaneDesc *)(host + o.planes);
  PlaneState *hState = (PlaneState *)(host + o.state);
  int4 *hRec = (int4 *)(host + o.rec);
  int *hCounters = (int *)(host + o.counters);
  for (int i = 0; i < n; i++) {
    DVC_CHECK(items[i].LR && items[i].acc && items[i].decoded, DVC_EINVAL, "dvc_ldpca: NULL item pointer");
    const int st0 = items[i].start_level;
    DVC_CHECK(st0 < C, DVC_EINVAL, "dvc_ldpca: start_level beyond the ladder");
    memcpy(host + o.lr + sizeof(float) * i * N, items[i].LR, sizeof(float) * N);
    memcpy(host + o.acc + (size_t)i * nb, items[i].acc, nb);
    if (mode1) memcpy(host + o.carry + (size_t)i * N, items[i].decoded, N);
    hPlanes[i] = PlaneDesc{N, dLR + (size_t)i * N, dLLR + (size_t)i * N, dAcc + (size_t)i * nb,
                           (uint32_t)items[i].crc, dState + i, dRec + (size_t)i * C,
                           mode1 ? dCarry + (size_t)i * N : nullptr, mode1 ? dBcase + (size_t)i * C : nullptr};
    hState[i] = PlaneState{st0 >= 0 ? st0 : -1, C - 1, 0.0f, 0, 0};
    for (int l = 0; l < C; l++) hRec[(size_t)i * C + l] = make_int4(0, REC_NOT_RUN, -1, 0);
  }
  hCounters[0] = hCounters[1] = hCounters[2] = 0;
  DVC_HIP(hipMemcpyAsync(dev, host, o.inEnd, hipMemcpyHostToDevice, s));
  TIMED(c, "ce_llr", s, launch_ce_llr(dPlanes, n, 1, c->lad1584, c->lad6336, s));
  std::vector<PlaneState> state(n);
  if (sequential) {
    DVC_HIP(hipMemcpyAsync(state.data(), dState, sizeof(PlaneState) * n, hipMemcpyDeviceToHost, s));
    DVC_HIP(hipStreamSynchronize(s));
  }

  const int threads = bp_threads(c, N);
  LadderDev ld = ladder_dev(*L);
  std::vector<int4> rec((size_t)n * C);
  if (sequential) {
    // Plain path: bitplane by bitplane, one code per launch, each with the
    // decisions the previous code left (ParallelMode 1).
    for (int i = 0; i < n; i++) {
      int level = state[i].start;
      int4 r;
      for (; level < C - 1;
           level = c->rate_step ? std::min(C - 1, state[i].start + dynstep_next(level - state[i].start)) : level + 1) {
        PlaneState ps{state[i].start, C - 1, state[i].entropy, level, state[i].taint};
        int zero[3] = {0, 0, 0};
        DVC_HIP(hipMemcpyAsync(dState + i, &ps, sizeof(ps), hipMemcpyHostToDevice, s));
        DVC_HIP(hipMemcpyAsync(dCounters, zero, sizeof(zero), hipMemcpyHostToDevice, s));
        BpArgs a{ld, dPlanes + i, 1, dCounters, 1, 1, c->max_iter, dSlots, dCounters + 1, slotCap, c->bp_poll,
                 c->bp_mode, c->rate_step};
        TIMED(c, N == 6336 ? "bp_6336" : "bp_1584", s, launch_bp_any(N, threads, a, 1, s));
        DVC_HIP(hipMemcpyAsync(&r, dRec + (size_t)i * C + level, sizeof(r), hipMemcpyDeviceToHost, s));
        DVC_HIP(hipStreamSynchronize(s));
        if (r.y == REC_ACCEPTED) {
          DVC_HIP(hipMemcpy(items[i].decoded, dSlots + (size_t)r.z * N, N, hipMemcpyDeviceToHost));
          break;
        }
        if (mode1) {
          DVC_CHECK(r.z >= 0 && r.z < slotCap, DVC_ESTATE, "dvc_ldpca: missing decisions of a failed code");
          DVC_HIP(hipMemcpyAsync(dCarry + (size_t)i * N, dSlots + (size_t)r.z * N, N, hipMemcpyDeviceToDevice, s));
        }
      }
      if (level == C - 1) {
        PlaneState ps{state[i].start, C - 1, state[i].entropy, C, state[i].taint};
        int zero[3] = {0, 0, 0};
        DVC_HIP(hipMemcpyAsync(dState + i, &ps, sizeof(ps), hipMemcpyHostToDevice, s));
        DVC_HIP(hipMemcpyAsync(dCounters, zero, sizeof(zero), hipMemcpyHostToDevice, s));
        TIMED(c, "invh", s, launch_invh(dPlanes + i, 1, ld, dSlots, dCounters + 1, slotCap, s));
        DVC_HIP(hipMemcpyAsync(&r, dRec + (size_t)i * C + (C - 1), sizeof(r), hipMemcpyDeviceToHost, s));
        DVC_HIP(hipStreamSynchronize(s));
        DVC_HIP(hipMemcpy(items[i].decoded, dSlots + (size_t)r.z * N, N, hipMemcpyDeviceToHost));
      }
    }
    DVC_HIP(hipMemcpy(rec.data(), dRec, sizeof(int4) * rec.size(), hipMemcpyDeviceToHost));
  } else {
    const int maxOff = C - 1;
    BpArgs a{ld, dPlanes, n, dCounters, n * maxOff, maxOff, c->max_iter, dSlots, dCounters + 1, slotCap,
             c->bp_poll, c->bp_mode, c->rate_step};
    const int grid = bp_grid(c, N, n * maxOff);
    TIMED(c, N == 6336 ? "bp_6336" : "bp_1584", s, launch_bp_any(N, threads, a, grid, s));
    for (int round = 0; mode1; round++) {
      // ParallelMode 1: resolve with the carried decisions; resume the
      // bitplanes that reach a code this launch did not finish
      DVC_CHECK(round <= C, DVC_ESTATE, "dvc_ldpca: ParallelMode 1 resolution does not settle");
      DVC_HIP(launch_zero_ints(dCounters + 2, 1, s));
      DVC_HIP(launch_mode1_resolve(dPlanes, n, C, N, dSlots, slotCap, dCounters + 2, s));
      int resume = 0;
      DVC_HIP(hipMemcpyAsync(&resume, dCounters + 2, sizeof(int), hipMemcpyDeviceToHost, s));
      DVC_HIP(hipStreamSynchronize(s));
      DVC_CHECK(resume < (1 << 20), DVC_ENOMEM, "ParallelMode 1: decision slots exhausted");
    if (!resume) break;
      DVC_HIP(launch_zero_ints(dCounters, 1, s));
      TIMED(c, N == 6336 ? "bp_6336" : "bp_1584", s, launch_bp_any(N, threads, a, grid, s));
    }
    TIMED(c, "invh", s, launch_invh(dPlanes, n, ld, dSlots, dCounters + 1, slotCap, s));
    // every bitplane's accepted decisions to dOut, then states, records,
    // counters and decisions back in one copy
    DVC_HIP(launch_gather_accepted(dPlanes, n, N, dSlots, slotCap, dOut, s));
    DVC_HIP(hipMemcpyAsync(host + o.state, dev + o.state, o.outEnd - o.state, hipMemcpyDeviceToHost, s));
    DVC_HIP(hipStreamSynchronize(s));
    DVC_CHECK(hCounters[1] <= slotCap, DVC_ENOMEM, "dvc_ldpca_decode_batch: decision slots exhausted");
    memcpy(state.data(), hState, sizeof(PlaneState) * n);
    memcpy(rec.data(), hRec, sizeof(int4) * rec.size());
    for (int i = 0; i < n; i++) {
      const int best = state[i].best;
      DVC_CHECK(best >= 0 && best < C, DVC_ESTATE, "dvc_ldpca_decode_batch: bad accepted code");
      const int4 r = rec[(size_t)i * C + best];
      DVC_CHECK(r.y == REC_ACCEPTED && r.z >= 0 && r.z < slotCap, DVC_ESTATE,
                "dvc_ldpca_decode_batch: missing accepted record");
      memcpy(items[i].decoded, host + o.out + (size_t)i * N, N);
    }
  }
  // Stats as the sequential search reports them.
  for (int i = 0; i < n; i++) {
    dvc_bp_stats &st = items[i].stats;
    memset(&st, 0, sizeof(st));
    st.start_level = state[i].start;
    st.entropy = state[i].entropy;
    int lvl = -1;
    for (int l = st.start_level; l < C; l++) {
      const int4 r = rec[(size_t)i * C + l];
      if (r.y == REC_ACCEPTED) { lvl = l; break; }
      st.iterations += r.x;
    }
    DVC_CHECK(lvl >= 0, DVC_ESTATE, "dvc_ldpca: no accepted code");
    st.level = lvl;
    st.iterations += rec[(size_t)i * C + lvl].x;
    st.rate = L->syndLength[lvl] + 8;
    st.requests = lvl - st.start_level + 1;
    if (c->rate_step) {
      // codes tried: the DYNAMICSTEP offsets below lvl, then lvl itself
      st.requests = 1;
      for (int off = 0; st.start_level + off < lvl; off = dynstep_next(off)) st.requests++;
    }
    st.used_inverse = rec[(size_t)i * C + lvl].w;
    for (int l = 0; l < C; l++) {
      const int4 r = rec[(size_t)i * C + l];
      if (r.y != REC_NOT_RUN) st.iterations_run += r.x;
    }
  }
  return DVC_OK;
}

int dvc_ldpca_decode(dvc_ctx *c, int N, const float *LR, const uint8_t *acc, uint8_t crc, int start_level,
                     uint8_t *decoded, dvc_bp_stats *stats) {
  DVC_CHECK(c, DVC_EINVAL, "dvc_ldpca_decode: NULL context");
  DVC_CHECK(c->ladder(N), DVC_EINVAL, "dvc_ldpca_decode: N must be 1584 or 6336");
  DVC_HIP(hipSetDevice(c->device));
  dvc_ldpca_item it{N, LR, acc, crc, start_level, decoded, {}};
  int rc = decode_group(c, N, 1, &it, 1, c->stream);
  if (rc == DVC_OK && stats) *stats = it.stats;
  return rc;
}

int dvc_ldpca_decode_batch(dvc_ctx *c, int n, dvc_ldpca_item *items) {
  DVC_CHECK(c, DVC_EINVAL, "dvc_ldpca_decode_batch: NULL context");
  DVC_CHECK(n >= 0 && (n == 0 || items), DVC_EINVAL, "dvc_ldpca_decode_batch: bad item list");
  DVC_HIP(hipSetDevice(c->device));
  std::vector<dvc_ldpca_item> g[2];
  std::vector<int> gi[2];
  for (int i = 0; i < n; i++) {
    DVC_CHECK(items[i].N == 1584 || items[i].N == 6336, DVC_EINVAL, "dvc_ldpca_decode_batch: N must be 1584 or 6336");
    int k = items[i].N == 6336;
    g[k].push_back(items[i]);
    gi[k].push_back(i);
  }
  for (int k = 0; k < 2; k++) {
    if (g[k].empty()) continue;
    int rc = decode_group(c, k ? 6336 : 1584, (int)g[k].size(), g[k].data(), 0, c->stream);
    if (rc) return rc;
    for (size_t j = 0; j < g[k].size(); j++) items[gi[k][j]].stats = g[k][j].stats;
  }
  return DVC_OK;
}

}  // extern "C"
