// frame_decoder.cpp -- host orchestration of WZ frame decoding on the GPU.
//
// The per-frame loop is decodeGOP's body (DvcDec/src/DVC_Decoder.cpp:26-474)
// with skip/intra/deblocking off: side information, residue transform and
// noise parameters, over-complete transform, then for each zig-zag band with
// bits: noise distribution; for each bit from MSB: conditional bit
// probability, conditional entropy, LDPCA rate search, bitplane insertion;
// then reconstruction, inverse transform, SI refresh, over-complete
// transform and motion learning; finally the inverse transform.
//
// All frame state stays on the device.  A batch of frames advances in
// lockstep, one bitplane per step; each kernel launch covers every frame that
// needs it in that step (frames of different sizes and quantisers mix freely,
// each following its own band schedule).  The host only decides which frames
// start or finish a band, and reads back, once per finished band, the rare
// reconstruction entries that lie within a rounding margin (recomputed with
// the C library's exp so the integers equal the CPU path's).
#include <math.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <vector>
#include "dvc_internal.h"
#include "frame_kernels.h"
#include "ldpca_kernels.h"

namespace dvc {

static const int kQTable[9][16] = {
    {0},
    {16, 8, 0, 0, 8, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {32, 8, 0, 0, 8, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {32, 8, 4, 0, 8, 4, 0, 0, 4, 0, 0, 0, 0, 0, 0, 0},
    {32, 16, 8, 4, 16, 8, 4, 0, 8, 4, 0, 0, 4, 0, 0, 0},
    {32, 16, 8, 4, 16, 8, 4, 4, 8, 4, 4, 0, 4, 4, 0, 0},
    {64, 16, 8, 8, 16, 8, 8, 4, 8, 8, 4, 4, 8, 4, 4, 0},
    {64, 32, 16, 8, 32, 16, 8, 4, 16, 8, 4, 4, 8, 4, 4, 0},
    {128, 64, 32, 16, 64, 32, 16, 8, 32, 16, 8, 4, 16, 8, 4, 0}};
static const int kQBits[9][16] = {
    {0},
    {4, 3, 0, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {5, 3, 0, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {5, 3, 2, 0, 3, 2, 0, 0, 2, 0, 0, 0, 0, 0, 0, 0},
    {5, 4, 3, 2, 4, 3, 2, 0, 3, 2, 0, 0, 2, 0, 0, 0},
    {5, 4, 3, 2, 4, 3, 2, 2, 3, 2, 2, 0, 2, 2, 0, 0},
    {6, 4, 3, 3, 4, 3, 3, 2, 3, 3, 2, 2, 3, 2, 2, 0},
    {6, 5, 4, 3, 5, 4, 3, 2, 4, 3, 2, 2, 3, 2, 2, 0},
    {7, 6, 5, 4, 6, 5, 4, 3, 5, 4, 3, 2, 4, 3, 2, 0}};
static const int kZigzag[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};

#define ROUNDD(x) ((x) >= 0.0 ? (int)((x) + 0.5) : -(int)(-(x) + 0.5))

struct BandPlan {
  int band, index, qbits, range, step, levelShift, stride;
};

// Band schedule of one frame (DVC_Decoder.cpp:139-161).
static std::vector<BandPlan> band_plan(int q, const int *acRange, int *bitrate0) {
  std::vector<BandPlan> bands;
  int step[16], crange[16];
  int br = 2;  // skip-mode and intra-mode flag bits (:131)
  step[0] = 1024 / kQTable[q][0];
  crange[0] = 1024;
  for (int i = 1; i < 16; i++) {
    if (kQTable[q][i] > 0) {
      step[i] = (int)ceil((double)((acRange[i] + 1) << 1) / (double)kQTable[q][i]);
      br += 10;
      crange[i] = ((kQTable[q][i] - 2) / 2 + 1) * step[i] - 1;
    } else {
      step[i] = 0;
      crange[i] = 0;
    }
  }
  for (int b = 0; b < 16; b++) {
    int idx = kZigzag[b];
    if (kQBits[q][idx] <= 0) break;
    BandPlan p;
    p.band = b; p.index = idx; p.qbits = kQBits[q][idx];
    p.range = crange[idx]; p.step = step[idx];
    p.levelShift = idx ? (kQTable[q][idx] - 2) / 2 : 0;
    p.stride = idx ? 2 * p.range + 1 : p.range;
    bands.push_back(p);
  }
  if (bitrate0) *bitrate0 = br;
  return bands;
}

struct KeyFrameDev {
  uint8_t *full = nullptr, *lp = nullptr, *half = nullptr;
};

struct Job {
  const dvc_wz_frame *spec = nullptr;
  int W = 0, H = 0, N = 0, nbp = 0;
  std::vector<BandPlan> bands;
  int bitrate0 = 0;
  FrameDev fd{};
  void *mem = nullptr;         // one device allocation for all per-frame buffers
  uint8_t *d_acc = nullptr;
  PlaneState *d_state = nullptr;
  int4 *d_rec = nullptr;
  float *d_LLR = nullptr;
  // cursor
  int bandPos = 0, bit = 0, bp = 0;
  bool done() const { return bandPos >= (int)bands.size(); }
};

struct Arena {
  std::vector<void *> ptrs;
  ~Arena() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t alloc(T **p, size_t count) {
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
};

static size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

// Carves the per-frame buffers out of one allocation.
static int alloc_job(Job &j, Arena &arena) {
  const int W = j.W, H = j.H, N = j.N, FS = W * H;
  int maxStride = 0;
  for (auto &b : j.bands) maxStride = std::max(maxStride, b.stride);
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off = align_up(off + bytes); return o; };
  const size_t o_si = take(FS), o_res = take(4 * (size_t)FS), o_st = take(4 * (size_t)FS), o_rt = take(4 * (size_t)FS),
               o_al = take(4 * (size_t)FS), o_oct = take((size_t)64 * (W - 3) * (H - 3)), o_smf = take((size_t)4 * 81 * N),
               o_dist = take((size_t)4 * N * std::max(maxStride, 1)), o_lr = take(4 * (size_t)N), o_dec = take(N),
               o_rec = take(FS), o_mb = take(sizeof(MVd) * (size_t)(FS / 256)), o_bmv = take(sizeof(MVd) * (size_t)(FS / 64)),
               o_bmvt = take(sizeof(MVd) * (size_t)(FS / 64)), o_fwd = take(sizeof(MVd) * (size_t)(FS / 256)),
               o_holes = take(16 * (size_t)(FS / 256)), o_mse = take(8 * (size_t)(FS / 64)), o_tc = take(4),
               o_tl = take(4 * (size_t)N), o_acc = take((size_t)j.nbp * (N / 8)), o_state = take(sizeof(PlaneState) * (size_t)j.nbp),
               o_recd = take(sizeof(int4) * 65 * (size_t)j.nbp), o_llr = take(4 * (size_t)N);
  uint8_t *base = nullptr;
  DVC_HIP(arena.alloc(&base, off));
  j.mem = base;
  FrameDev &f = j.fd;
  f.si = base + o_si; f.residue = (int *)(base + o_res); f.sideTrans = (int *)(base + o_st);
  f.reconsTrans = (int *)(base + o_rt); f.alpha = (float *)(base + o_al); f.oct = (int *)(base + o_oct);
  f.smf = (float *)(base + o_smf); f.dist = (float *)(base + o_dist); f.LR = (float *)(base + o_lr);
  f.decoded = base + o_dec; f.recons = base + o_rec; f.mbmv = (MVd *)(base + o_mb); f.bmv = (MVd *)(base + o_bmv);
  f.bmv_tmp = (MVd *)(base + o_bmvt); f.fwd = (MVd *)(base + o_fwd); f.holes = (double *)(base + o_holes);
  f.mse = (double *)(base + o_mse); f.tieCount = (int *)(base + o_tc); f.tieList = (int *)(base + o_tl);
  j.d_acc = base + o_acc; j.d_state = (PlaneState *)(base + o_state); j.d_rec = (int4 *)(base + o_recd);
  j.d_LLR = (float *)(base + o_llr);
  return DVC_OK;
}

// reconstruction (reconstruction.cpp:184-289) of one coefficient with the C
// library's exp, for entries the kernel left near a rounding boundary.
static int recon_host(int rt, int side, float alphaF, int index, int step, int ls) {
  int lo, hi;
  const double a = alphaF;
  if (index == 0) {
    lo = rt * step; hi = (rt + 1) * step;
  } else {
    const int q = rt - ls;
    if (q == 0) { lo = -step + 1; hi = step; }
    else if (q > 0) { lo = q * step; hi = (q + 1) * step; }
    else { lo = (q - 1) * step + 1; hi = q * step + 1; }
  }
  if (a <= 0.0001) return (lo + hi) / 2;
  if (side < lo) return lo + ROUNDD(1.0 / a + (double)step / (1.0 - exp(a * (double)step)));
  if (side >= hi) return hi - ROUNDD(1.0 / a + (double)step / (1.0 - exp(a * (double)step)));
  const double r = side - lo, d = hi - side;
  const double er = exp(-a * r), ed = exp(-a * d);
  return side + ROUNDD(((r + 1.0 / a) * er - (d + 1.0 / a) * ed) / (2.0 - (er + ed)));
}

}  // namespace dvc

using namespace dvc;

// ---------------------------------------------------------------- context extras
namespace dvc {
int ensure_cost_tables(dvc_ctx *c) {
  if (c->d_costFme) return DVC_OK;
  std::vector<double> fme(65 * 65), usi(65 * 65);
  const double K = (double)0.05f;
  for (int dy = -32; dy <= 32; dy++)
    for (int dx = -32; dx <= 32; dx++) {
      double x = dx, y = dy;
      fme[(dy + 32) * 65 + dx + 32] = 1.0 + K * sqrt(x * x + y * y);
      double x2 = (double)(dx << 1), y2 = (double)(dy << 1);
      usi[(dy + 32) * 65 + dx + 32] = 1.0 + K * sqrt(x2 * x2 + y2 * y2);
    }
  DVC_HIP(hipMalloc(&c->d_costFme, fme.size() * sizeof(double)));
  DVC_HIP(hipMalloc(&c->d_costUsi, usi.size() * sizeof(double)));
  DVC_HIP(hipMemcpy(c->d_costFme, fme.data(), fme.size() * sizeof(double), hipMemcpyHostToDevice));
  DVC_HIP(hipMemcpy(c->d_costUsi, usi.data(), usi.size() * sizeof(double), hipMemcpyHostToDevice));
  return DVC_OK;
}

#define TIMED_LAUNCH(c, name, s, expr)                           \
  do {                                                           \
    int slot_;                                                   \
    hipEvent_t ev_ = (c)->timer_begin(name, s, &slot_);          \
    DVC_HIP(expr);                                               \
    (c)->timer_end(slot_, ev_, s);                               \
  } while (0)

// One bitplane step of LDPCA decoding for the planes of one code length.
int bp_step(dvc_ctx *c, int N, const std::vector<PlaneDesc> &planes, bool sequential, uint8_t *d_slots, int slotCap,
            int *d_counters, PlaneDesc *d_planes, hipStream_t s);
}  // namespace dvc

int dvc::bp_step(dvc_ctx *c, int N, const std::vector<PlaneDesc> &planes, bool sequential, uint8_t *d_slots,
                 int slotCap, int *d_counters, PlaneDesc *d_planes, hipStream_t s) {
  Ladder *L = c->ladder(N);
  const int n = (int)planes.size(), C = L->numCodes;
  if (n == 0) return DVC_OK;
  LadderDev ld = ladder_dev(*L);
  const int threads = N == 1584 ? 256 : (c->bp_threads_cif > 0 ? c->bp_threads_cif : 1024);
  DVC_HIP(hipMemcpyAsync(d_planes, planes.data(), sizeof(PlaneDesc) * n, hipMemcpyHostToDevice, s));
  TIMED_LAUNCH(c, "ce_llr", s, launch_ce_llr(d_planes, n, 1, c->lad1584, c->lad6336, s));
  if (!sequential) {
    const int maxOff = C - 1;
    DVC_HIP(hipMemsetAsync(d_counters, 0, 2 * sizeof(int), s));
    BpArgs a{ld, d_planes, n, d_counters, n * maxOff, maxOff, c->max_iter, d_slots, d_counters + 1, slotCap};
    int grid = std::min(c->sm_count * (N == 6336 ? 1 : 4), n * maxOff);
    TIMED_LAUNCH(c, N == 6336 ? "bp_6336" : "bp_1584", s, launch_bp_any(N, threads, a, grid, s));
    TIMED_LAUNCH(c, "invh", s, launch_invh(d_planes, n, ld, d_slots, d_counters + 1, slotCap, s));
    return DVC_OK;
  }
  // Plain path: per bitplane, one code per launch in increasing order, with
  // the host deciding the next code (the reference's beliefPropagation_CUDA loop).
  // Every host-side source below stays alive until the stream is drained.
  std::vector<PlaneState> st(n);
  DVC_HIP(hipMemsetAsync(d_counters, 0, 2 * sizeof(int), s));
  for (int i = 0; i < n; i++)
    DVC_HIP(hipMemcpyAsync(&st[i], planes[i].state, sizeof(PlaneState), hipMemcpyDeviceToHost, s));
  DVC_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < n; i++) {
    int level = st[i].start;
    bool ok = false;
    PlaneState ps;
    int4 r;
    for (; level < C - 1; level++) {
      ps = PlaneState{level, C - 1, st[i].entropy, 0};
      DVC_HIP(hipMemcpyAsync(planes[i].state, &ps, sizeof(ps), hipMemcpyHostToDevice, s));
      DVC_HIP(hipMemsetAsync(d_counters, 0, sizeof(int), s));
      BpArgs a{ld, d_planes + i, 1, d_counters, 1, 1, c->max_iter, d_slots, d_counters + 1, slotCap};
      TIMED_LAUNCH(c, N == 6336 ? "bp_6336" : "bp_1584", s, launch_bp_any(N, threads, a, 1, s));
      DVC_HIP(hipMemcpyAsync(&r, planes[i].rec + level, sizeof(r), hipMemcpyDeviceToHost, s));
      DVC_HIP(hipStreamSynchronize(s));
      if (r.y == REC_ACCEPTED) { ok = true; break; }
    }
    ps = PlaneState{st[i].start, ok ? level : C - 1, st[i].entropy, 0};
    DVC_HIP(hipMemcpyAsync(planes[i].state, &ps, sizeof(ps), hipMemcpyHostToDevice, s));
    if (!ok) TIMED_LAUNCH(c, "invh", s, launch_invh(d_planes + i, 1, ld, d_slots, d_counters + 1, slotCap, s));
    DVC_HIP(hipStreamSynchronize(s));
  }
  return DVC_OK;
}

extern "C" int dvc_decode_wz_frames(dvc_ctx *c, int n, const dvc_wz_frame *frames, int batched) {
  DVC_CHECK(c, DVC_EINVAL, "dvc_decode_wz_frames: NULL context");
  DVC_CHECK(n >= 0 && (n == 0 || frames), DVC_EINVAL, "dvc_decode_wz_frames: bad frame list");
  if (n == 0) return DVC_OK;
  DVC_HIP(hipSetDevice(c->device));
  int rc = ensure_cost_tables(c);
  if (rc) return rc;
  hipStream_t s = c->stream;
  Arena arena;
  // ---- key frames: upload and filter each distinct host frame once
  std::map<const uint8_t *, KeyFrameDev> keys;
  std::vector<Job> jobs(n);
  for (int i = 0; i < n; i++) {
    const dvc_wz_frame &sp = frames[i];
    DVC_CHECK((sp.width == 176 && sp.height == 144) || (sp.width == 352 && sp.height == 288), DVC_EINVAL,
              "dvc_decode_wz_frames: frame size must be QCIF or CIF");
    DVC_CHECK(sp.qindex >= 1 && sp.qindex <= 8, DVC_EINVAL, "dvc_decode_wz_frames: qindex must be 1..8");
    DVC_CHECK(sp.past && sp.fut && sp.ac_range && sp.crc && sp.acc && sp.recons, DVC_EINVAL,
              "dvc_decode_wz_frames: NULL frame pointer");
    DVC_CHECK(sp.past_idx < sp.cur_idx && sp.cur_idx < sp.fut_idx, DVC_EINVAL,
              "dvc_decode_wz_frames: need past_idx < cur_idx < fut_idx");
    Job &j = jobs[i];
    j.spec = &sp;
    j.W = sp.width; j.H = sp.height; j.N = j.W * j.H / 16;
    j.bands = band_plan(sp.qindex, sp.ac_range, &j.bitrate0);
    for (auto &b : j.bands) j.nbp += b.qbits;
    for (const uint8_t *kp : {sp.past, sp.fut}) {
      if (keys.count(kp)) continue;
      KeyFrameDev k;
      const int FS = j.W * j.H;
      DVC_HIP(arena.alloc(&k.full, FS));
      DVC_HIP(arena.alloc(&k.lp, FS));
      DVC_HIP(arena.alloc(&k.half, 4 * (size_t)FS));
      DVC_HIP(hipMemcpyAsync(k.full, kp, FS, hipMemcpyHostToDevice, s));
      TIMED_LAUNCH(c, "keyframe_filters", s, launch_keyframe_filters(k.full, k.lp, k.half, j.W, j.H, s));
      keys[kp] = k;
    }
    rc = alloc_job(j, arena);
    if (rc) return rc;
    FrameDev &f = j.fd;
    f.W = j.W; f.H = j.H; f.N = j.N; f.qindex = sp.qindex;
    f.iPC = (double)(sp.cur_idx - sp.past_idx) / (double)(sp.fut_idx - sp.past_idx);
    f.iFC = (double)(sp.fut_idx - sp.cur_idx) / (double)(sp.fut_idx - sp.past_idx);
    const KeyFrameDev &kp = keys[sp.past], &kf = keys[sp.fut];
    f.past = kp.full; f.past_lp = kp.lp; f.past_half = kp.half;
    f.fut = kf.full; f.fut_lp = kf.lp; f.fut_half = kf.half;
    DVC_HIP(hipMemcpyAsync(j.d_acc, sp.acc, (size_t)j.nbp * (j.N / 8), hipMemcpyHostToDevice, s));
    DVC_HIP(hipMemsetAsync(f.reconsTrans, 0, sizeof(int) * j.W * j.H, s));
    std::vector<int4> recInit((size_t)65 * j.nbp, make_int4(0, REC_NOT_RUN, -1, 0));
    DVC_HIP(hipMemcpy(j.d_rec, recInit.data(), recInit.size() * sizeof(int4), hipMemcpyHostToDevice));
  }
  // ---- descriptors
  FrameDev *d_frames;
  int *d_lists;  // up to 4 lists of n entries
  DVC_HIP(arena.alloc(&d_frames, n));
  DVC_HIP(arena.alloc(&d_lists, 4 * (size_t)n));
  std::vector<FrameDev> hf(n);
  auto upload_frames = [&]() -> int {
    for (int i = 0; i < n; i++) hf[i] = jobs[i].fd;
    DVC_HIP(hipMemcpyAsync(d_frames, hf.data(), sizeof(FrameDev) * n, hipMemcpyHostToDevice, s));
    return DVC_OK;
  };
  std::vector<int> hlists(4 * (size_t)n);
  auto upload_list = [&](int slot, const std::vector<int> &l) -> int {
    std::copy(l.begin(), l.end(), hlists.begin() + (size_t)slot * n);
    DVC_HIP(hipMemcpyAsync(d_lists + (size_t)slot * n, l.data(), sizeof(int) * l.size(), hipMemcpyHostToDevice, s));
    return DVC_OK;
  };
  CostTables ct{c->d_costFme, c->d_costUsi};
  std::vector<int> all(n);
  int maxN = 0, maxMB = 0, maxPos = 0;
  for (int i = 0; i < n; i++) {
    all[i] = i;
    maxN = std::max(maxN, jobs[i].N);
    maxMB = std::max(maxMB, jobs[i].W * jobs[i].H / 256);
    maxPos = std::max(maxPos, (jobs[i].W - 3) * (jobs[i].H - 3));
  }
  if ((rc = upload_frames()) || (rc = upload_list(0, all))) return rc;
  // ---- side information, noise parameters, over-complete transform, SMF
  TIMED_LAUNCH(c, "create_si", s, launch_create_si(d_frames, d_lists, n, maxMB, ct, s));
  TIMED_LAUNCH(c, "noise_param", s, launch_noise_param(d_frames, d_lists, n, maxN, s));
  TIMED_LAUNCH(c, "oct", s, launch_oct(d_frames, d_lists, n, maxPos, 1, s));
  TIMED_LAUNCH(c, "smf_init", s, launch_smf_init(d_frames, d_lists, n, maxN, s));

  // ---- bitplane loop
  // every (bitplane, code) item accepts at most once, plus one full-rate slot:
  // 65 slots per frame and code-length group can never overflow
  const int groupCap = 65 * n + 1;
  int slotCap = 2 * groupCap;
  uint8_t *d_slots;
  int *d_counters;
  PlaneDesc *d_planes;
  BpFrameLink *d_links;
  DVC_HIP(arena.alloc(&d_slots, (size_t)slotCap * 6336));
  DVC_HIP(arena.alloc(&d_counters, 2));
  DVC_HIP(arena.alloc(&d_planes, n));
  DVC_HIP(arena.alloc(&d_links, n));
  for (auto &j : jobs) { j.bandPos = 0; j.bit = j.bands.empty() ? 0 : j.bands[0].qbits; j.bp = 0; }
  int host_fixups = 0;
  for (;;) {
    std::vector<int> active, starting, ending;
    for (int i = 0; i < n; i++) {
      Job &j = jobs[i];
      if (j.done()) continue;
      const BandPlan &b = j.bands[j.bandPos];
      if (j.bit == b.qbits) {
        starting.push_back(i);
        j.fd.band = b.band; j.fd.index = b.index; j.fd.range = b.range; j.fd.step = b.step;
        j.fd.levelShift = b.levelShift; j.fd.distStride = b.stride;
      }
      j.fd.bit = j.bit;
      active.push_back(i);
      if (j.bit == 1) ending.push_back(i);
    }
    if (active.empty()) break;
    if ((rc = upload_frames())) return rc;
    if (!starting.empty()) {
      int maxStride = 0;
      for (int i : starting) {
        maxStride = std::max(maxStride, jobs[i].fd.distStride);
        DVC_CHECK(jobs[i].fd.index == 0 || jobs[i].fd.range >= 0, DVC_EINVAL, "bad coefficient range");
      }
      if ((rc = upload_list(1, starting))) return rc;
      TIMED_LAUNCH(c, "noise_dist", s, launch_noise_dist(d_frames, d_lists + n, (int)starting.size(), maxN, maxStride, s));
    }
    if ((rc = upload_list(0, active))) return rc;
    TIMED_LAUNCH(c, "condbit", s, launch_condbit(d_frames, d_lists, (int)active.size(), maxN, s));
    // LDPCA for this bitplane of every active frame, grouped by code length
    std::vector<PlaneDesc> groups[2];
    std::vector<BpFrameLink> links(active.size());
    std::vector<PlaneState> init(active.size(), PlaneState{-1, 64, 0.0f, 0});
    for (size_t a = 0; a < active.size(); a++) {
      Job &j = jobs[active[a]];
      DVC_HIP(hipMemcpyAsync(j.d_state + j.bp, &init[a], sizeof(PlaneState), hipMemcpyHostToDevice, s));
      PlaneDesc pd{j.N, j.fd.LR, j.d_LLR, j.d_acc + (size_t)j.bp * (j.N / 8), (uint32_t)j.spec->crc[j.bp],
                   j.d_state + j.bp, j.d_rec + (size_t)65 * j.bp};
      groups[j.N == 6336].push_back(pd);
      links[a] = BpFrameLink{j.d_state + j.bp, j.d_rec + (size_t)65 * j.bp, d_slots, groupCap};
    }
    DVC_HIP(hipStreamSynchronize(s));
    int planeOff = 0;
    for (int g = 0; g < 2; g++) {
      if (groups[g].empty()) continue;
      // each group uses its own slice of the slot pool
      const int cap = groupCap;
      rc = bp_step(c, g ? 6336 : 1584, groups[g], !batched, d_slots + (size_t)(g ? cap : 0) * 6336, cap,
                   d_counters, d_planes + planeOff, s);
      if (rc) return rc;
      planeOff += (int)groups[g].size();
      DVC_HIP(hipStreamSynchronize(s));
    }
    // slots of group 1 start at cap * 6336 bytes; k_put_bitplane indexes slots by N, so give it the
    // group-specific base pointer and per-N stride through the link.
    for (size_t a = 0; a < active.size(); a++) {
      Job &j = jobs[active[a]];
      links[a].slots = d_slots + (size_t)(j.N == 6336 ? groupCap : 0) * 6336;
      links[a].slot_capacity = groupCap;
    }
    DVC_HIP(hipMemcpyAsync(d_links, links.data(), sizeof(BpFrameLink) * links.size(), hipMemcpyHostToDevice, s));
    TIMED_LAUNCH(c, "put_bitplane", s, launch_put_bitplane(d_frames, d_lists, d_links, (int)active.size(), maxN, s));
    for (int i : active) {
      Job &j = jobs[i];
      j.bp++;
      j.bit--;
    }
    if (!ending.empty()) {
      if ((rc = upload_list(2, ending))) return rc;
      const int *le = d_lists + 2 * (size_t)n;
      const int ne = (int)ending.size();
      for (int i : ending) DVC_HIP(hipMemsetAsync(jobs[i].fd.tieCount, 0, sizeof(int), s));
      TIMED_LAUNCH(c, "reconstruct", s, launch_reconstruct(d_frames, le, ne, maxN, s));
      // entries near a rounding boundary: recompute with the C library's exp
      std::vector<int> ties(ne);
      for (int k = 0; k < ne; k++)
        DVC_HIP(hipMemcpyAsync(&ties[k], jobs[ending[k]].fd.tieCount, sizeof(int), hipMemcpyDeviceToHost, s));
      DVC_HIP(hipStreamSynchronize(s));
      for (int k = 0; k < ne; k++) {
        Job &j = jobs[ending[k]];
        const int cnt = std::min(ties[k], j.N);
        DVC_CHECK(ties[k] <= j.N, DVC_ESTATE, "reconstruction tie list overflow");
        if (cnt == 0) continue;
        std::vector<int> idx(cnt);
        DVC_HIP(hipMemcpy(idx.data(), j.fd.tieList, sizeof(int) * cnt, hipMemcpyDeviceToHost));
        for (int ci : idx) {
          int rt, side;
          float al;
          DVC_HIP(hipMemcpy(&rt, j.fd.reconsTrans + ci, sizeof(int), hipMemcpyDeviceToHost));
          DVC_HIP(hipMemcpy(&side, j.fd.sideTrans + ci, sizeof(int), hipMemcpyDeviceToHost));
          DVC_HIP(hipMemcpy(&al, j.fd.alpha + ci, sizeof(float), hipMemcpyDeviceToHost));
          int v = recon_host(rt, side, al, j.fd.index, j.fd.step, j.fd.levelShift);
          DVC_HIP(hipMemcpy(j.fd.sideTrans + ci, &v, sizeof(int), hipMemcpyHostToDevice));
          host_fixups++;
        }
      }
      TIMED_LAUNCH(c, "inverse_transform", s, launch_inverse_transform(d_frames, le, ne, maxN, s));
      TIMED_LAUNCH(c, "update_si", s, launch_update_si(d_frames, le, ne, maxN, ct, s));
      TIMED_LAUNCH(c, "oct", s, launch_oct(d_frames, le, ne, maxPos, 0, s));
      TIMED_LAUNCH(c, "update_smf", s, launch_update_smf(d_frames, le, ne, maxN, s));
      TIMED_LAUNCH(c, "update_si_ml", s, launch_update_si_ml(d_frames, le, ne, maxN, 0, s));
      for (int i : ending) {
        Job &j = jobs[i];
        j.bandPos++;
        if (!j.done()) j.bit = j.bands[j.bandPos].qbits;
      }
    }
    // host buffers of this step (lists, links, descriptors) are reused next step
    DVC_HIP(hipStreamSynchronize(s));
  }
  // ---- final inverse transform (DVC_Decoder.cpp:411-423), outputs and statistics
  if ((rc = upload_list(3, all))) return rc;
  TIMED_LAUNCH(c, "inverse_transform", s, launch_inverse_transform(d_frames, d_lists + 3 * (size_t)n, n, maxN, s));
  for (int i = 0; i < n; i++) {
    Job &j = jobs[i];
    DVC_HIP(hipMemcpyAsync(j.spec->recons, j.fd.recons, j.W * j.H, hipMemcpyDeviceToHost, s));
  }
  DVC_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < n; i++) {
    Job &j = jobs[i];
    dvc_frame_stats *st = j.spec->stats;
    if (!st) continue;
    memset(st, 0, sizeof(*st));
    st->nbitplanes = j.nbp;
    st->bitrate = j.bitrate0;
    std::vector<PlaneState> ps(j.nbp);
    std::vector<int4> rec((size_t)65 * j.nbp);
    DVC_HIP(hipMemcpy(ps.data(), j.d_state, sizeof(PlaneState) * j.nbp, hipMemcpyDeviceToHost));
    DVC_HIP(hipMemcpy(rec.data(), j.d_rec, sizeof(int4) * rec.size(), hipMemcpyDeviceToHost));
    Ladder *L = c->ladder(j.N);
    for (int b = 0; b < j.nbp && b < 128; b++) {
      const int4 *r = &rec[(size_t)65 * b];
      int lvl = -1, it = 0;
      for (int l = ps[b].start; l < 65; l++) {
        if (r[l].y == REC_ACCEPTED) { lvl = l; break; }
        it += r[l].x;
      }
      DVC_CHECK(lvl >= 0, DVC_ESTATE, "dvc_decode_wz_frames: bitplane without accepted code");
      st->levels[b] = lvl;
      st->start_levels[b] = ps[b].start;
      st->iterations[b] = it + r[lvl].x;
      st->used_inverse[b] = r[lvl].w;
      st->entropy[b] = ps[b].entropy;
      st->bitrate += L->syndLength[lvl] + 8;
    }
    st->recon_host_fixups = host_fixups;
    if (j.spec->orig) {
      double mse = 0.0;
      const int FS = j.W * j.H;
      for (int p = 0; p < FS; p++) {
        const double d = (double)j.spec->recons[p] - (double)j.spec->orig[p];
        mse += d * d;
      }
      mse /= (double)FS;
      st->psnr = mse == 0.0 ? 0.0 : 10.0 * log10(255.0 * 255.0 / mse);  // frame.cpp:115-124
    }
  }
  return DVC_OK;
}
