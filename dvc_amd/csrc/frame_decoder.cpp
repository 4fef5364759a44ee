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
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <tuple>
#include <vector>
#include "dvc_internal.h"
#include "frame_kernels.h"
#include "ldpca_kernels.h"
#include "engine_kernels.h"

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

int qtable_entry(int q, int index) { return (q >= 1 && q <= 8 && index >= 0 && index < 16) ? kQTable[q][index] : 0; }
int qbits_entry(int q, int index) { return (q >= 1 && q <= 8 && index >= 0 && index < 16) ? kQBits[q][index] : 0; }
int zigzag_entry(int band) { return (band >= 0 && band < 16) ? kZigzag[band] : 0; }

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
  dvc_wz_frame spec{};           // caller's request; pointers must stay valid until completion
  std::vector<int> acRange;      // owned copies of the bitstream fields
  std::vector<uint8_t> crc, acc;
  std::vector<uint8_t> mode;     // 8x8 block modes when the frame has skip or intra blocks, else empty
  std::vector<int32_t> intra;    // intra blocks' dequantised coefficients (16 per 4x4 block), else empty
  std::vector<uint8_t> pastMode; // deblocking: the past reference's block modes (a WZ frame's), else empty
  bool deblock = false;
  uint64_t ticket = 0, pastKey = 0, futKey = 0;
  bool setup = false;
  size_t bytes = 0;
  int W = 0, H = 0, N = 0, nbp = 0;
  std::vector<BandPlan> bands;
  int bitrate0 = 0;
  FrameDev fd{};
  void *mem = nullptr;         // one device allocation for all per-frame buffers
  uint8_t *d_acc = nullptr;
  PlaneState *d_state = nullptr;
  int4 *d_rec = nullptr;
  float *d_LLR = nullptr;
  uint8_t *d_carry = nullptr;  // ParallelMode 1: decisions carried from code to code and bitplane to bitplane
  int *d_bcase = nullptr;      // ParallelMode 1: [nbp][kMaxCodes] repeat-case decision slots
  // cursor
  int bandPos = 0, bit = 0, bp = 0;
  bool done() const { return bandPos >= (int)bands.size(); }
};

static size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

// Device bytes one frame needs (the layout alloc_job carves).
static size_t job_bytes(const Job &j) {
  const int W = j.W, H = j.H, N = j.N, FS = W * H;
  int maxStride = 0;
  for (auto &b : j.bands) maxStride = std::max(maxStride, b.stride);
  size_t t = 0;
  for (size_t b : {(size_t)FS, 4 * (size_t)FS, 4 * (size_t)FS, 4 * (size_t)FS, 4 * (size_t)FS,
                   (size_t)64 * (W - 3) * (H - 3), (size_t)4 * (W - 3) * (H - 3), (size_t)4 * 81 * N, (size_t)4 * N * std::max(maxStride, 1),
                   4 * (size_t)N, (size_t)N, (size_t)FS, sizeof(MVd) * (size_t)(FS / 256), sizeof(MVd) * (size_t)(FS / 64),
                   sizeof(MVd) * (size_t)(FS / 64), sizeof(MVd) * (size_t)(FS / 256), 16 * (size_t)(FS / 256),
                   8 * (size_t)(FS / 64), (size_t)j.nbp * (N / 8),
                   sizeof(PlaneState) * (size_t)j.nbp, sizeof(int4) * kMaxCodes * (size_t)j.nbp, 4 * (size_t)N,
                   (size_t)N, 4 * kMaxCodes * (size_t)j.nbp, (size_t)FS / 64, j.intra.size() * sizeof(int32_t),
                   j.deblock ? (size_t)FS : 0, j.deblock ? 2 * (size_t)N : 0, j.deblock ? 16 * (size_t)N : 0,
                   j.pastMode.size()})
    t = align_up(t + b);
  return t;
}

// Carves the per-frame buffers out of one block of j.bytes (base).
static void carve_job(Job &j, uint8_t *base) {
  const int W = j.W, H = j.H, N = j.N, FS = W * H;
  int maxStride = 0;
  for (auto &b : j.bands) maxStride = std::max(maxStride, b.stride);
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off = align_up(off + bytes); return o; };
  const size_t o_si = take(FS), o_res = take(4 * (size_t)FS), o_st = take(4 * (size_t)FS), o_rt = take(4 * (size_t)FS),
               o_al = take(4 * (size_t)FS), o_oct = take((size_t)64 * (W - 3) * (H - 3)),
               o_ob = take((size_t)4 * (W - 3) * (H - 3)), o_smf = take((size_t)4 * 81 * N),
               o_dist = take((size_t)4 * N * std::max(maxStride, 1)), o_lr = take(4 * (size_t)N), o_dec = take(N),
               o_rec = take(FS), o_mb = take(sizeof(MVd) * (size_t)(FS / 256)), o_bmv = take(sizeof(MVd) * (size_t)(FS / 64)),
               o_bmvt = take(sizeof(MVd) * (size_t)(FS / 64)), o_fwd = take(sizeof(MVd) * (size_t)(FS / 256)),
               o_holes = take(16 * (size_t)(FS / 256)), o_mse = take(8 * (size_t)(FS / 64)),
               o_acc = take((size_t)j.nbp * (N / 8)), o_state = take(sizeof(PlaneState) * (size_t)j.nbp),
               o_recd = take(sizeof(int4) * kMaxCodes * (size_t)j.nbp), o_llr = take(4 * (size_t)N),
               o_carry = take(N), o_bcase = take(4 * kMaxCodes * (size_t)j.nbp), o_mode = take((size_t)FS / 64),
               o_intra = take(j.intra.size() * sizeof(int32_t)), o_dbf = take(j.deblock ? (size_t)FS : 0),
               o_dbm = take(j.deblock ? 2 * (size_t)N : 0), o_dbb = take(j.deblock ? 16 * (size_t)N : 0),
               o_pm = take(j.pastMode.size());
  j.mem = base;
  j.bytes = off;
  FrameDev &f = j.fd;
  f.si = base + o_si; f.residue = (int *)(base + o_res); f.sideTrans = (int *)(base + o_st);
  f.reconsTrans = (int *)(base + o_rt); f.alpha = (float *)(base + o_al); f.oct = (int *)(base + o_oct);
  f.octBand = (int *)(base + o_ob);
  f.smf = (float *)(base + o_smf); f.dist = (float *)(base + o_dist); f.LR = (float *)(base + o_lr);
  f.decoded = base + o_dec; f.recons = base + o_rec; f.mbmv = (MVd *)(base + o_mb); f.bmv = (MVd *)(base + o_bmv);
  f.bmv_tmp = (MVd *)(base + o_bmvt); f.fwd = (MVd *)(base + o_fwd); f.holes = (double *)(base + o_holes);
  f.mse = (double *)(base + o_mse);
  j.d_acc = base + o_acc; j.d_state = (PlaneState *)(base + o_state); j.d_rec = (int4 *)(base + o_recd);
  j.d_LLR = (float *)(base + o_llr);
  j.d_carry = base + o_carry;
  j.d_bcase = (int *)(base + o_bcase);
  f.mode = j.mode.empty() ? nullptr : base + o_mode;
  f.intraCoeffs = j.intra.empty() ? nullptr : (int *)(base + o_intra);
  f.dbFlag = j.deblock ? base + o_dbf : nullptr;
  f.dbEdgeMode = j.deblock ? base + o_dbm : nullptr;
  f.dbEdgeBs = j.deblock ? (double *)(base + o_dbb) : nullptr;
  f.pastMode = j.pastMode.empty() ? nullptr : base + o_pm;
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
// The speculative step's launches alone (the plane descriptors already in
// d_planes): entropy and LLR, counters, the persistent BP launch and the
// full-rate inverse.  Kernel launches only, so a stream capture of it replays.
int bp_step_launch(dvc_ctx *c, int N, int n, uint8_t *d_slots, int slotCap, int *d_counters, PlaneDesc *d_planes,
                   hipStream_t s);
}  // namespace dvc

int dvc::bp_step_launch(dvc_ctx *c, int N, int n, uint8_t *d_slots, int slotCap, int *d_counters, PlaneDesc *d_planes,
                        hipStream_t s) {
  Ladder *L = c->ladder(N);
  const int C = L->numCodes;
  LadderDev ld = ladder_dev(*L);
  const int threads = bp_threads(c, N);
  const int maxOff = C - 1;
  TIMED_LAUNCH(c, "ce_llr", s, launch_ce_llr(d_planes, n, 1, c->lad1584, c->lad6336, s));
  DVC_HIP(launch_zero_ints(d_counters, 3, s));
  BpArgs a{ld, d_planes, n, d_counters, n * maxOff, maxOff, c->max_iter, d_slots, d_counters + 1, slotCap, c->bp_poll,
           c->bp_mode, c->rate_step};
  const int grid = bp_grid(c, N, n * maxOff);
  TIMED_LAUNCH(c, N == 6336 ? "bp_6336" : "bp_1584", s, launch_bp_any(N, threads, a, grid, s));
  for (int round = 0; c->bp_mode == 1; round++) {
    // ParallelMode 1: resolve with the carried decisions and resume the
    // bitplanes that reach a code the launch did not finish (host-synchronous)
    DVC_CHECK(round <= C, DVC_ESTATE, "ParallelMode 1 resolution does not settle");
    DVC_HIP(launch_zero_ints(d_counters + 2, 1, s));
    DVC_HIP(launch_mode1_resolve(d_planes, n, C, N, d_slots, slotCap, d_counters + 2, s));
    int resume = 0;
    DVC_HIP(hipMemcpyAsync(&resume, d_counters + 2, sizeof(int), hipMemcpyDeviceToHost, s));
    DVC_HIP(hipStreamSynchronize(s));
    DVC_CHECK(resume < (1 << 20), DVC_ENOMEM, "ParallelMode 1: decision slots exhausted");
    if (!resume) break;
    DVC_HIP(launch_zero_ints(d_counters, 1, s));
    TIMED_LAUNCH(c, N == 6336 ? "bp_6336" : "bp_1584", s, launch_bp_any(N, threads, a, grid, s));
  }
  TIMED_LAUNCH(c, "invh", s, launch_invh(d_planes, n, ld, d_slots, d_counters + 1, slotCap, s));
  return DVC_OK;
}

int dvc::bp_step(dvc_ctx *c, int N, const std::vector<PlaneDesc> &planes, bool sequential, uint8_t *d_slots,
                 int slotCap, int *d_counters, PlaneDesc *d_planes, hipStream_t s) {
  Ladder *L = c->ladder(N);
  const int n = (int)planes.size(), C = L->numCodes;
  if (n == 0) return DVC_OK;
  LadderDev ld = ladder_dev(*L);
  const int threads = bp_threads(c, N);
  DVC_HIP(hipMemcpyAsync(d_planes, planes.data(), sizeof(PlaneDesc) * n, hipMemcpyHostToDevice, s));
  if (!sequential) return bp_step_launch(c, N, n, d_slots, slotCap, d_counters, d_planes, s);
  TIMED_LAUNCH(c, "ce_llr", s, launch_ce_llr(d_planes, n, 1, c->lad1584, c->lad6336, s));
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
    for (; level < C - 1;
         level = c->rate_step ? std::min(C - 1, st[i].start + dynstep_next(level - st[i].start)) : level + 1) {
      ps = PlaneState{st[i].start, C - 1, st[i].entropy, level, st[i].taint};
      DVC_HIP(hipMemcpyAsync(planes[i].state, &ps, sizeof(ps), hipMemcpyHostToDevice, s));
      DVC_HIP(hipMemsetAsync(d_counters, 0, sizeof(int), s));  // queue head only: the step's slots accumulate
      BpArgs a{ld, d_planes + i, 1, d_counters, 1, 1, c->max_iter, d_slots, d_counters + 1, slotCap, c->bp_poll,
               c->bp_mode, c->rate_step};
      TIMED_LAUNCH(c, N == 6336 ? "bp_6336" : "bp_1584", s, launch_bp_any(N, threads, a, 1, s));
      DVC_HIP(hipMemcpyAsync(&r, planes[i].rec + level, sizeof(r), hipMemcpyDeviceToHost, s));
      DVC_HIP(hipStreamSynchronize(s));
      if (r.y == REC_ACCEPTED) { ok = true; break; }
      if (c->bp_mode == 1) {
        // the failed code's decisions are carried into the next code
        DVC_CHECK(r.z >= 0 && r.z < slotCap, DVC_ESTATE, "missing decisions of a failed code");
        DVC_HIP(hipMemcpyAsync(planes[i].carry, d_slots + (size_t)r.z * N, N, hipMemcpyDeviceToDevice, s));
      }
    }
    ps = PlaneState{st[i].start, ok ? level : C - 1, st[i].entropy, C, st[i].taint};
    DVC_HIP(hipMemcpyAsync(planes[i].state, &ps, sizeof(ps), hipMemcpyHostToDevice, s));
    if (!ok) TIMED_LAUNCH(c, "invh", s, launch_invh(d_planes + i, 1, ld, d_slots, d_counters + 1, slotCap, s));
    DVC_HIP(hipStreamSynchronize(s));
  }
  return DVC_OK;
}


// ---------------------------------------------------------------- engine
namespace dvc {
struct KeyEntry {
  KeyFrameDev k;
  int W = 0, H = 0;
  int refs = 0;
  bool named = false;     // kept resident after its last use (a stream key)
  uint64_t lastUse = 0;
  size_t bytes = 0;
};

// One lane of the engine: a stream, the frames it advances in lockstep and
// its device scratch.  Lanes step concurrently from host threads, so the tail
// of one lane's LDPCA launch and its host round trips overlap the other
// lanes' kernels.
struct Lane {
  hipStream_t s = nullptr;
  int cap = 0, groupCap = 0;
  std::vector<Job *> inflight;
  std::vector<Job *> finished;  // retired by the engine after the step
  // A step's descriptors, written to one pinned host block and uploaded to
  // its device mirror with one copy: frame descriptors, five frame lists of
  // cap entries (stepped, starting a band, ending a band, set up, finishing),
  // BP frame links, plane descriptors, result copy segments (3 per frame).
  uint8_t *h_stage = nullptr, *d_stage = nullptr;
  size_t stageBytes = 0;
  FrameDev *h_frames = nullptr, *d_frames = nullptr;
  int *h_lists = nullptr, *d_lists = nullptr;
  BpFrameLink *h_links = nullptr, *d_links = nullptr;
  PlaneDesc *h_planes = nullptr, *d_planes = nullptr;
  CopySeg *h_segs = nullptr, *d_segs = nullptr;
  // results of the step's finishing frames (rate-search states, run records,
  // reconstructions the host needs), gathered on the device and read back
  // with one copy; grown on demand
  uint8_t *h_out = nullptr, *d_out = nullptr;
  size_t outBytes = 0;
  uint8_t *d_slots = nullptr;
  int *d_counters = nullptr;
  int rc = 0;
  std::string err;
  // host-side lists of one step
  std::vector<int> l0, l1, l2, l3, l4;
  std::vector<PlaneDesc> groups[2];
  // Captured launch sequences of steady steps (no frame starts or ends a
  // band), keyed by their launch shape: frames stepped, planes of each code
  // length, largest code length.  Their kernels read the step's frame list,
  // frame descriptors and plane descriptors from this lane's buffers, which
  // are uploaded before each replay.
  std::map<std::tuple<int, int, int, int>, hipGraphExec_t> graphs;
};
}  // namespace dvc

struct dvc_engine {
  dvc_ctx *c = nullptr;
  size_t budget = 0, used = 0;
  int maxInflight = 1;
  bool sequentialBp = false;
  bool graphs = true;  // replay captured steady steps
  bool deviceBuffers = false;  // past, fut, acc and recons of requests are device pointers
  int laneSteps = 0;           // steps a lane may take per engine step (dvc_engine_params.lane_steps)
  uint64_t nextTicket = 1, clock = 0;
  std::deque<dvc::Job *> pending;
  std::vector<dvc::Lane> lanes;
  std::map<uint64_t, int> doneTickets;  // ticket -> 1
  std::map<uint64_t, dvc::KeyEntry> keys;
  // device blocks of retired frames and dropped key frames, kept for reuse
  // by size (counted in `pooled`, not in `used`)
  std::multimap<size_t, void *> pool;
  size_t pooled = 0;
  // admission descriptors: frame initialisations then key frames, one
  // pinned block and its device mirror
  uint8_t *h_admit = nullptr, *d_admit = nullptr;
  size_t admitBytes = 0, admitKeysOff = 0;
  int inflight() const {
    int n = 0;
    for (auto &l : lanes) n += (int)l.inflight.size();
    return n;
  }
};

namespace dvc {

// Device block of `bytes` from the pool (a retired block of that size) or a
// new allocation.
static int pool_alloc(dvc_engine *e, size_t bytes, void **p) {
  auto it = e->pool.find(bytes);
  if (it != e->pool.end()) {
    *p = it->second;
    e->pool.erase(it);
    e->pooled -= bytes;
    return DVC_OK;
  }
  DVC_HIP(hipMalloc(p, bytes));
  return DVC_OK;
}

// Keeps a block whose last user has finished (its stream was synchronised).
static void pool_free(dvc_engine *e, void *p, size_t bytes) {
  if (!p) return;
  e->pool.emplace(bytes, p);
  e->pooled += bytes;
}

// Frees pooled blocks while live, pooled and needed bytes exceed the budget.
static void pool_trim(dvc_engine *e, size_t need) {
  while (!e->pool.empty() && e->used + e->pooled + need > e->budget) {
    auto it = std::prev(e->pool.end());
    (void)hipFree(it->second);
    e->pooled -= it->first;
    e->pool.erase(it);
  }
}

static void free_job(dvc_engine *e, Job *j) {
  pool_free(e, j->mem, j->bytes);
  j->mem = nullptr;
  e->used -= j->bytes;
  for (uint64_t k : {j->pastKey, j->futKey}) {
    auto it = e->keys.find(k);
    if (it == e->keys.end()) continue;
    it->second.refs--;
  }
  delete j;
}

// Drops unreferenced anonymous key frames, then the least recently used named
// ones while the budget is exceeded.
static void trim_keys(dvc_engine *e, size_t need) {
  for (auto it = e->keys.begin(); it != e->keys.end();) {
    if (it->second.refs == 0 && !it->second.named) {
      KeyFrameDev &k = it->second.k;
      pool_free(e, k.full, it->second.bytes);
      e->used -= it->second.bytes;
      it = e->keys.erase(it);
    } else {
      ++it;
    }
  }
  while (e->used + need > e->budget) {
    auto victim = e->keys.end();
    for (auto it = e->keys.begin(); it != e->keys.end(); ++it)
      if (it->second.refs == 0 && (victim == e->keys.end() || it->second.lastUse < victim->second.lastUse)) victim = it;
    if (victim == e->keys.end()) break;
    pool_free(e, victim->second.k.full, victim->second.bytes);
    e->used -= victim->second.bytes;
    e->keys.erase(victim);
  }
}

// A key frame for a frame being admitted: the resident one under `key`, or
// a new one whose filtering is queued in `jobs` (device requests are read in
// place; host requests are copied to the device first).
static int key_frame(dvc_engine *e, uint64_t key, bool named, const uint8_t *src, int W, int H, hipStream_t s,
                     KeyFrameDev *out, std::vector<KeyFrameJob> &jobs) {
  auto it = e->keys.find(key);
  if (it != e->keys.end()) {
    DVC_CHECK(it->second.W == W && it->second.H == H, DVC_EINVAL, "key frame reused with another size");
    it->second.refs++;
    it->second.lastUse = ++e->clock;
    *out = it->second.k;
    return DVC_OK;
  }
  KeyEntry ke;
  const size_t FS = (size_t)W * H;
  ke.bytes = 6 * FS;
  uint8_t *p = nullptr;
  int rc = pool_alloc(e, ke.bytes, (void **)&p);
  if (rc) return rc;
  ke.k.full = p; ke.k.lp = p + FS; ke.k.half = p + 2 * FS;
  ke.W = W; ke.H = H; ke.refs = 1; ke.named = named; ke.lastUse = ++e->clock;
  if (!e->deviceBuffers) DVC_HIP(hipMemcpyAsync(ke.k.full, src, FS, hipMemcpyHostToDevice, s));
  jobs.push_back(KeyFrameJob{e->deviceBuffers ? src : ke.k.full, ke.k.full, ke.k.lp, ke.k.half, W, H});
  e->used += ke.bytes;
  e->keys[key] = ke;
  *out = ke.k;
  return DVC_OK;
}

// Gives a frame its device state and queues its initialisation (`inits`)
// and new key frames (`keyJobs`) for the admission's batched launches on
// stream s; the caller synchronises s before any lane uses the frame.
static int admit(dvc_engine *e, Job *j, hipStream_t s, std::vector<JobInitDev> &inits,
                 std::vector<KeyFrameJob> &keyJobs) {
  void *mem = nullptr;
  int rc = pool_alloc(e, j->bytes, &mem);
  if (rc) return rc;
  carve_job(*j, (uint8_t *)mem);
  e->used += j->bytes;
  FrameDev &f = j->fd;
  const dvc_wz_frame &sp = j->spec;
  f.W = j->W; f.H = j->H; f.N = j->N; f.qindex = sp.qindex;
  f.iPC = (double)(sp.cur_idx - sp.past_idx) / (double)(sp.fut_idx - sp.past_idx);
  f.iFC = (double)(sp.fut_idx - sp.cur_idx) / (double)(sp.fut_idx - sp.past_idx);
  KeyFrameDev kp, kf;
  const bool pn = j->pastKey != (uint64_t)(uintptr_t)sp.past, fn = j->futKey != (uint64_t)(uintptr_t)sp.fut;
  if ((rc = key_frame(e, j->pastKey, pn, sp.past, j->W, j->H, s, &kp, keyJobs))) return rc;
  if ((rc = key_frame(e, j->futKey, fn, sp.fut, j->W, j->H, s, &kf, keyJobs))) return rc;
  f.past = kp.full; f.past_lp = kp.lp; f.past_half = kp.half;
  f.fut = kf.full; f.fut_lp = kf.lp; f.fut_half = kf.half;
  if (!e->deviceBuffers) DVC_HIP(hipMemcpyAsync(j->d_acc, j->acc.data(), j->acc.size(), hipMemcpyHostToDevice, s));
  if (f.mode) DVC_HIP(hipMemcpyAsync((void *)f.mode, j->mode.data(), j->mode.size(), hipMemcpyHostToDevice, s));
  if (f.intraCoeffs)
    DVC_HIP(hipMemcpyAsync((void *)f.intraCoeffs, j->intra.data(), j->intra.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  if (f.pastMode)
    DVC_HIP(hipMemcpyAsync((void *)f.pastMode, j->pastMode.data(), j->pastMode.size(), hipMemcpyHostToDevice, s));
  f.deblock = j->deblock ? 1 : 0;
  if (j->deblock) {
    // deblock.cpp:361-365 and 56-58: coefficient weights Qtable[i] / sum, the
    // filter's QP (fixed at its first call to the stream's IntraQP) and 2^(QP/6)
    int sum = 0;
    for (int i = 0; i < 16; i++) sum += qtable_entry(sp.qindex, i);
    for (int i = 0; i < 16; i++) f.dbWeight[i] = (double)qtable_entry(sp.qindex, i) / (double)sum;
    f.intraQp = sp.intra_qp;
    f.pow2qp = pow(2, ((double)sp.intra_qp / 6.0));
    f.pastIsWz = sp.past_is_wz ? 1 : 0;
  }
  // k_job_init: no decoded bits; per-bitplane rate-search state with the
  // start level from the entropy (computed on the device), no accepted code
  // and no record of any code run; ParallelMode 1's carried decisions 0xAA
  // (the reference reads its per-frame decoded buffer before anything is
  // written to it, DVC_Decoder.cpp:124, and in its golden runs that buffer
  // holds the MALLOC_PERTURB_ fill 0xAA, which never matches a decision)
  inits.push_back(JobInitDev{f.reconsTrans, j->W * j->H, j->d_carry, j->N, j->d_state, j->d_rec, j->nbp,
                             e->c->ladder(j->N)->numCodes, e->deviceBuffers ? j->spec.acc : nullptr, j->d_acc,
                             j->nbp * (j->N / 8)});
  j->bandPos = 0;
  j->bit = j->bands.empty() ? 0 : j->bands[0].qbits;
  j->bp = 0;
  return DVC_OK;
}

// Bytes of a finishing frame's gathered results: rate-search states, run
// records and, when the host needs it, the reconstruction.
static bool job_needs_host_recons(const dvc_engine *e, const Job &j) {
  return !e->deviceBuffers || (j.spec.orig && j.spec.stats);
}
static size_t job_out_bytes(const dvc_engine *e, const Job &j) {
  return align_up(sizeof(PlaneState) * j.nbp) + align_up(sizeof(int4) * kMaxCodes * (size_t)j.nbp) +
         (job_needs_host_recons(e, j) ? align_up((size_t)j.W * j.H) : 0);
}

// A finished frame's outputs from its gathered results `out` (host copy, in
// the layout of job_out_bytes; a device request's reconstruction was copied
// to the caller's buffer on the device).
// Every bitplane's accepted record must name a decision slot below the
// lane's capacity (slot_cap): k_put_bitplane skips a bitplane whose slot lies
// beyond it, so such a frame is reported as DVC_ENOMEM instead of finishing.
static int finish_job(dvc_engine *e, Job &j, const uint8_t *out, int slot_cap) {
  dvc_ctx *c = e->c;
  const size_t FS = (size_t)j.W * j.H;
  const PlaneState *ps = (const PlaneState *)out;
  const int4 *rec = (const int4 *)(out + align_up(sizeof(PlaneState) * j.nbp));
  const uint8_t *rec8 = job_needs_host_recons(e, j)
                            ? out + align_up(sizeof(PlaneState) * j.nbp) + align_up(sizeof(int4) * kMaxCodes * (size_t)j.nbp)
                            : nullptr;
  const int C = c->ladder(j.N)->numCodes;
  for (int b = 0; b < j.nbp; b++) {
    const int best = ps[b].best;
    DVC_CHECK(best >= 0 && best < C, DVC_ESTATE, "WZ frame decode: bitplane without accepted code");
    const int4 r = rec[(size_t)kMaxCodes * b + best];
    DVC_CHECK(r.y == REC_ACCEPTED, DVC_ESTATE, "WZ frame decode: bitplane without accepted code");
    DVC_CHECK(r.z >= 0 && r.z < slot_cap, DVC_ENOMEM,
              "WZ frame decode: decision slots exhausted (an accepted bitplane's decisions were not stored)");
  }
  if (!e->deviceBuffers) memcpy(j.spec.recons, rec8, FS);
  dvc_frame_stats *st = j.spec.stats;
  if (!st) return DVC_OK;
  memset(st, 0, sizeof(*st));
  st->nbitplanes = j.nbp;
  st->bitrate = j.bitrate0;
  Ladder *L = c->ladder(j.N);
  for (int b = 0; b < j.nbp && b < 128; b++) {
    const int4 *r = &rec[(size_t)kMaxCodes * b];
    int lvl = -1, it = 0;
    for (int l = std::max(ps[b].start, 0); l < C; l++) {
      if (r[l].y == REC_ACCEPTED) { lvl = l; break; }
      it += r[l].x;
    }
    DVC_CHECK(lvl >= 0, DVC_ESTATE, "WZ frame decode: bitplane without accepted code");
    st->levels[b] = lvl;
    st->start_levels[b] = ps[b].start;
    st->iterations[b] = it + r[lvl].x;
    st->used_inverse[b] = r[lvl].w;
    st->entropy[b] = ps[b].entropy;
    st->bitrate += L->syndLength[lvl] + 8;
  }
  if (j.spec.orig) {
    double mse = 0.0;
    for (size_t p = 0; p < FS; p++) {
      const double d = (double)rec8[p] - (double)j.spec.orig[p];
      mse += d * d;
    }
    mse /= (double)FS;
    st->psnr = mse == 0.0 ? 0.0 : 10.0 * log10(255.0 * 255.0 / mse);  // frame.cpp:115-124
  }
  return DVC_OK;
}

// One bitplane of every frame of a lane (with the band starts and ends it
// implies), setting up the lane's new frames first and finishing the frames
// whose last band ended.  Runs on the lane's stream only.
static int lane_step(dvc_engine *e, Lane &ln) {
  dvc_ctx *c = e->c;
  hipStream_t s = ln.s;
  int rc;
  const int n = (int)ln.inflight.size();
  if (n == 0) return DVC_OK;
  int maxN = 0, maxMB = 0, maxPos = 0, maxBw = 0, maxBh = 0;
  for (int i = 0; i < n; i++) {
    Job *j = ln.inflight[i];
    maxN = std::max(maxN, j->N);
    maxMB = std::max(maxMB, j->W * j->H / 256);
    maxPos = std::max(maxPos, (j->W - 3) * (j->H - 3));
    maxBw = std::max(maxBw, j->W / 4);
    maxBh = std::max(maxBh, j->H / 4);
  }
  CostTables ct{c->d_costFme, c->d_costUsi};
  // ---- the step's frame lists (all known before any kernel runs)
  std::vector<int> &active = ln.l0, &starting = ln.l1, &ending = ln.l2, &setup = ln.l3, &fin = ln.l4;
  for (auto *l : {&active, &starting, &ending, &setup, &fin}) l->clear();
  for (int i = 0; i < n; i++)
    if (!ln.inflight[i]->setup) setup.push_back(i);
  for (int i = 0; i < n; i++) {
    Job &j = *ln.inflight[i];
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
  // frames whose last band ends in this step (or that have none)
  for (int i = 0; i < n; i++) {
    const Job &j = *ln.inflight[i];
    if (j.done() || (j.bit == 1 && j.bandPos + 1 == (int)j.bands.size())) fin.push_back(i);
  }
  const int M = ln.cap;
  const int *L0 = ln.d_lists, *L1 = ln.d_lists + M, *L2 = ln.d_lists + 2 * M, *L3 = ln.d_lists + 3 * M,
            *L4 = ln.d_lists + 4 * M;
  {
    int k = 0;
    for (auto *l : {&active, &starting, &ending, &setup, &fin}) {
      if (!l->empty()) memcpy(ln.h_lists + (size_t)k * M, l->data(), sizeof(int) * l->size());
      k++;
    }
  }
  for (int i = 0; i < n; i++) ln.h_frames[i] = ln.inflight[i]->fd;
  // ---- LDPCA descriptors of the stepped frames
  ln.groups[0].clear();
  ln.groups[1].clear();
  const bool m1 = c->bp_mode == 1;
  for (size_t a = 0; a < active.size(); a++) {
    Job &j = *ln.inflight[active[a]];
    PlaneDesc pd{j.N, j.fd.LR, j.d_LLR, j.d_acc + (size_t)j.bp * (j.N / 8), (uint32_t)j.crc[j.bp],
                 j.d_state + j.bp, j.d_rec + (size_t)kMaxCodes * j.bp, m1 ? j.d_carry : nullptr,
                 m1 ? j.d_bcase + (size_t)kMaxCodes * j.bp : nullptr};
    ln.groups[j.N == 6336].push_back(pd);
    ln.h_links[a] = BpFrameLink{j.d_state + j.bp, j.d_rec + (size_t)kMaxCodes * j.bp,
                                ln.d_slots + (size_t)(j.N == 6336 ? ln.groupCap : 0) * 6336, ln.groupCap,
                                m1 ? j.d_carry : nullptr};
  }
  const int na = (int)active.size(), n0 = (int)ln.groups[0].size(), n1 = (int)ln.groups[1].size();
  if (n0) memcpy(ln.h_planes, ln.groups[0].data(), sizeof(PlaneDesc) * n0);
  if (n1) memcpy(ln.h_planes + n0, ln.groups[1].data(), sizeof(PlaneDesc) * n1);
  // ---- result gathering of the finishing frames: copy segments into the
  // lane's output block (and, for device requests, to the caller's buffer)
  size_t outUsed = 0;
  std::vector<size_t> outOff(fin.size());
  for (size_t k = 0; k < fin.size(); k++) {
    outOff[k] = outUsed;
    outUsed += job_out_bytes(e, *ln.inflight[fin[k]]);
  }
  if (outUsed > ln.outBytes) {
    // the lane's stream is idle here (synchronised at the end of every step)
    if (ln.d_out) (void)hipFree(ln.d_out);
    if (ln.h_out) (void)hipHostFree(ln.h_out);
    ln.d_out = ln.h_out = nullptr;
    ln.outBytes = 0;
    const size_t want = outUsed + outUsed / 4;
    DVC_HIP(hipMalloc((void **)&ln.d_out, want));
    DVC_HIP(hipHostMalloc((void **)&ln.h_out, want, hipHostMallocDefault));
    ln.outBytes = want;
  }
  int nseg = 0;
  for (size_t k = 0; k < fin.size(); k++) {
    const Job &j = *ln.inflight[fin[k]];
    uint8_t *o = ln.d_out + outOff[k];
    const size_t bs = sizeof(PlaneState) * j.nbp, br = sizeof(int4) * kMaxCodes * (size_t)j.nbp, FS = (size_t)j.W * j.H;
    ln.h_segs[nseg++] = CopySeg{(const uint8_t *)j.d_state, o, bs};
    ln.h_segs[nseg++] = CopySeg{(const uint8_t *)j.d_rec, o + align_up(bs), br};
    if (job_needs_host_recons(e, j)) ln.h_segs[nseg++] = CopySeg{j.fd.recons, o + align_up(bs) + align_up(br), FS};
    if (e->deviceBuffers) ln.h_segs[nseg++] = CopySeg{j.fd.recons, j.spec.recons, FS};
  }
  // ---- one upload of the step's descriptors
  DVC_HIP(hipMemcpyAsync(ln.d_stage, ln.h_stage, ln.stageBytes, hipMemcpyHostToDevice, s));
  // ---- setup of newly admitted frames: SI, noise parameters, OCT, SMF
  if (!setup.empty()) {
    const int nf = (int)setup.size();
    TIMED_LAUNCH(c, "create_si", s, launch_create_si(ln.d_frames, L3, nf, maxMB, ct, s));
    // intra blocks (intraRecon, DVC_Decoder.cpp:43-47) and their copy into the
    // side information (:56), before the over-complete transform reads it
    bool anyIntra = false;
    for (int i : setup) anyIntra |= !ln.inflight[i]->intra.empty();
    if (anyIntra) TIMED_LAUNCH(c, "intra", s, launch_intra(ln.d_frames, L3, nf, maxN / 4, s));
    TIMED_LAUNCH(c, "noise_param", s, launch_noise_param(ln.d_frames, L3, nf, maxN, s));
    TIMED_LAUNCH(c, "oct", s, launch_oct(ln.d_frames, L3, nf, maxPos, 1, s));
    TIMED_LAUNCH(c, "smf_init", s, launch_smf_init(ln.d_frames, L3, nf, maxN, s));
    for (int i : setup) ln.inflight[i]->setup = true;
  }
  if (!active.empty()) {
    if (!starting.empty()) {
      int maxStride = 0;
      for (int i : starting) maxStride = std::max(maxStride, ln.inflight[i]->fd.distStride);
      TIMED_LAUNCH(c, "noise_dist", s, launch_noise_dist(ln.d_frames, L1, (int)starting.size(), maxN, maxStride, s));
    }
    // the step's kernels: conditional bit probabilities, LDPCA per code
    // length, bitplane insertion
    auto launches = [&]() -> int {
      TIMED_LAUNCH(c, "condbit", s, launch_condbit(ln.d_frames, L0, na, maxN, s));
      int off = 0;
      for (int g = 0; g < 2; g++) {
        const int ng = g ? n1 : n0;
        if (!ng) continue;
        uint8_t *slots = ln.d_slots + (size_t)(g ? ln.groupCap : 0) * 6336;
        const int r = e->sequentialBp
                          ? bp_step(c, g ? 6336 : 1584, ln.groups[g], true, slots, ln.groupCap, ln.d_counters + 3 * g,
                                    ln.d_planes + off, s)
                          : bp_step_launch(c, g ? 6336 : 1584, ng, slots, ln.groupCap, ln.d_counters + 3 * g,
                                           ln.d_planes + off, s);
        if (r) return r;
        off += ng;
      }
      TIMED_LAUNCH(c, "put_bitplane", s, launch_put_bitplane(ln.d_frames, L0, ln.d_links, na, maxN, s));
      return DVC_OK;
    };
    // (ParallelMode 1 steps read a resolution count back, so they are not captured)
    const bool steady = e->graphs && !e->sequentialBp && !c->profiling && c->bp_mode == 0 && setup.empty() &&
                        starting.empty() && ending.empty();
    if (steady) {
      const auto key = std::make_tuple(na, n0, n1, maxN);
      auto it = ln.graphs.find(key);
      if (it == ln.graphs.end()) {
        DVC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int lrc = launches();
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(s, &graph);
        if (lrc) {
          if (graph) (void)hipGraphDestroy(graph);
          return lrc;
        }
        DVC_HIP(ce);
        hipGraphExec_t exec = nullptr;
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        DVC_HIP(ie);
        it = ln.graphs.emplace(key, exec).first;
      }
      DVC_HIP(hipGraphLaunch(it->second, s));
    } else if ((rc = launches())) {
      return rc;
    }
    for (int i : active) {
      ln.inflight[i]->bp++;
      ln.inflight[i]->bit--;
    }
    if (!ending.empty()) {
      const int ne = (int)ending.size();
      TIMED_LAUNCH(c, "reconstruct", s, launch_reconstruct(ln.d_frames, L2, ne, maxN, s));
      TIMED_LAUNCH(c, "inverse_transform", s, launch_inverse_transform(ln.d_frames, L2, ne, maxN, s));
      TIMED_LAUNCH(c, "update_si", s, launch_update_si(ln.d_frames, L2, ne, maxN, ct, s));
      TIMED_LAUNCH(c, "oct", s, launch_oct(ln.d_frames, L2, ne, maxPos, 0, s));
      TIMED_LAUNCH(c, "update_smf", s, launch_update_smf(ln.d_frames, L2, ne, maxN, s));
      TIMED_LAUNCH(c, "update_si_ml", s, launch_update_si_ml(ln.d_frames, L2, ne, maxBw, maxBh, s));
      for (int i : ending) {
        Job &j = *ln.inflight[i];
        j.bandPos++;
        if (!j.done()) j.bit = j.bands[j.bandPos].qbits;
      }
    }
  }
  // ---- retire: final inverse transform (DVC_Decoder.cpp:411-423), results
  // gathered on the device and read back with one copy, then outputs
  if (!fin.empty()) {
    TIMED_LAUNCH(c, "inverse_transform", s, launch_inverse_transform(ln.d_frames, L4, (int)fin.size(), maxN, s));
    // skip blocks take the past key frame's pixels (DVC_Decoder.cpp:155-157);
    // no kernel of the band loop reads or writes them, so they are copied here
    bool anySkip = false;
    for (int i : fin) anySkip |= !ln.inflight[i]->mode.empty();
    if (anySkip) DVC_HIP(launch_skip_copy(ln.d_frames, L4, (int)fin.size(), maxN / 4, s));
    // the deblocking filter (DVC_Decoder.cpp:429-435) on the finished frame
    bool anyDeblock = false;
    int dw = 0, dh = 0;
    for (int i : fin) {
      const Job &j = *ln.inflight[i];
      if (!j.deblock) continue;
      anyDeblock = true;
      dw = std::max(dw, j.W);
      dh = std::max(dh, j.H);
    }
    if (anyDeblock) TIMED_LAUNCH(c, "deblock", s, launch_deblock(ln.d_frames, L4, (int)fin.size(), dw, dh, s));
    DVC_HIP(launch_copy_segments(ln.d_segs, nseg, s));
    DVC_HIP(hipMemcpyAsync(ln.h_out, ln.d_out, outUsed, hipMemcpyDeviceToHost, s));
    DVC_HIP(hipStreamSynchronize(s));
    for (size_t k = 0; k < fin.size(); k++) {
      Job &j = *ln.inflight[fin[k]];
      DVC_CHECK(j.setup && j.done(), DVC_ESTATE, "engine: a finishing frame has bands left");
      if ((rc = finish_job(e, j, ln.h_out + outOff[k], ln.groupCap))) return rc;
    }
    std::vector<Job *> keep;
    for (int i = 0; i < n; i++) {
      Job *j = ln.inflight[i];
      if (j->setup && j->done()) ln.finished.push_back(j);
      else keep.push_back(j);
    }
    ln.inflight.swap(keep);
  }
  DVC_HIP(hipStreamSynchronize(s));  // the step's host staging is reused next step
  return DVC_OK;
}

// One engine step: admit queued frames while the HBM budget and the lanes'
// capacity allow (uploads on the context stream, finished before any lane
// runs), step every lane with frames (concurrently when several have), then
// retire the lanes' finished frames.
static int engine_step(dvc_engine *e, int *completed) {
  dvc_ctx *c = e->c;
  int rc;
  *completed = 0;
  bool admitted = false;
  std::vector<JobInitDev> inits;
  std::vector<KeyFrameJob> keyJobs;
  while (!e->pending.empty() && e->inflight() < e->maxInflight) {
    Job *j = e->pending.front();
    Lane *dst = nullptr;
    for (auto &l : e->lanes)
      if ((int)l.inflight.size() < l.cap && (!dst || l.inflight.size() < dst->inflight.size())) dst = &l;
    if (!dst) break;
    const size_t keyNeed = 12 * (size_t)j->W * j->H;
    trim_keys(e, j->bytes + keyNeed);
    pool_trim(e, j->bytes + keyNeed);
    if (e->inflight() > 0 && e->used + j->bytes + keyNeed > e->budget) break;
    e->pending.pop_front();
    if ((rc = admit(e, j, c->stream, inits, keyJobs))) { delete j; return rc; }
    dst->inflight.push_back(j);
    admitted = true;
  }
  if (admitted) {
    // one upload of the admission's descriptors, then the batched key-frame
    // filters and frame initialisation
    const size_t bi = sizeof(JobInitDev) * inits.size(), bk = sizeof(KeyFrameJob) * keyJobs.size();
    DVC_CHECK(bi <= e->admitKeysOff && e->admitKeysOff + bk <= e->admitBytes, DVC_ESTATE,
              "engine: admission staging too small");
    if (bi) memcpy(e->h_admit, inits.data(), bi);
    if (bk) memcpy(e->h_admit + e->admitKeysOff, keyJobs.data(), bk);
    DVC_HIP(hipMemcpyAsync(e->d_admit, e->h_admit, e->admitBytes, hipMemcpyHostToDevice, c->stream));
    int maxW = 0, maxH = 0;
    for (const KeyFrameJob &k : keyJobs) { maxW = std::max(maxW, k.W); maxH = std::max(maxH, k.H); }
    TIMED_LAUNCH(c, "keyframe_filters", c->stream,
                 launch_keyframe_batch((const KeyFrameJob *)(e->d_admit + e->admitKeysOff), (int)keyJobs.size(), maxW,
                                       maxH, c->stream));
    DVC_HIP(launch_job_init((const JobInitDev *)e->d_admit, (int)inits.size(), c->stream));
    DVC_HIP(hipStreamSynchronize(c->stream));
  }
  std::vector<Lane *> busy;
  for (auto &l : e->lanes)
    if (!l.inflight.empty()) busy.push_back(&l);
  if (busy.empty()) return DVC_OK;
  // a lane keeps stepping without waiting for the others until one of its
  // frames completes or lane_steps steps are done; one step while frames
  // wait for admission (the pending queue cannot change during the call)
  const int maxSteps = e->pending.empty() ? std::max(1, e->laneSteps) : 1;
  auto run = [e, c, maxSteps](Lane *l) {
    l->rc = DVC_OK;
    if (hipSetDevice(c->device) != hipSuccess) { l->rc = DVC_EDEVICE; l->err = "hipSetDevice failed"; return; }
    for (int k = 0; k < maxSteps; k++) {
      l->rc = lane_step(e, *l);
      if (l->rc) { l->err = dvc_last_error(); return; }
      if (!l->finished.empty() || l->inflight.empty()) return;
    }
  };
  std::vector<std::thread> th;
  for (size_t k = 1; k < busy.size(); k++) th.emplace_back(run, busy[k]);
  run(busy[0]);
  for (auto &t : th) t.join();
  for (Lane *l : busy)
    if (l->rc) { set_error(l->err); return l->rc; }
  for (auto &l : e->lanes) {
    for (Job *j : l.finished) {
      e->doneTickets[j->ticket] = 1;
      free_job(e, j);
      (*completed)++;
    }
    l.finished.clear();
  }
  return DVC_OK;
}

static int engine_submit(dvc_engine *e, const dvc_wz_frame *sp, uint64_t pastKey, uint64_t futKey, uint64_t *ticket) {
  DVC_CHECK(sp, DVC_EINVAL, "dvc_engine_submit: NULL frame");
  DVC_CHECK((sp->width == 176 && sp->height == 144) || (sp->width == 352 && sp->height == 288), DVC_EINVAL,
            "WZ frame size must be QCIF or CIF");
  DVC_CHECK(sp->qindex >= 1 && sp->qindex <= 8, DVC_EINVAL, "qindex must be 1..8");
  DVC_CHECK(sp->past && sp->fut && sp->ac_range && sp->crc && sp->acc && sp->recons, DVC_EINVAL,
            "WZ frame: NULL pointer");
  DVC_CHECK(sp->past_idx < sp->cur_idx && sp->cur_idx < sp->fut_idx, DVC_EINVAL, "need past_idx < cur_idx < fut_idx");
  Job *j = new Job();
  j->spec = *sp;
  j->W = sp->width; j->H = sp->height; j->N = j->W * j->H / 16;
  j->acRange.assign(sp->ac_range, sp->ac_range + 16);
  j->bands = band_plan(sp->qindex, j->acRange.data(), &j->bitrate0);
  const int n8 = j->W * j->H / 64;
  if (sp->block_mode) {
    bool anyNonWz = false, anyIntra = false;
    for (int k = 0; k < n8; k++) {
      DVC_CHECK(sp->block_mode[k] <= 2, DVC_EINVAL, "WZ frame: block modes must be 0 (WZ), 1 (intra) or 2 (skip)");
      anyNonWz |= sp->block_mode[k] != 0;
      anyIntra |= sp->block_mode[k] == 1;
    }
    if (anyNonWz) j->mode.assign(sp->block_mode, sp->block_mode + n8);
    if (anyIntra) {
      DVC_CHECK(sp->intra_coeffs, DVC_EINVAL, "WZ frame: intra blocks need their coefficients (intra_coeffs)");
      j->intra.assign(sp->intra_coeffs, sp->intra_coeffs + (size_t)16 * j->N);
    }
  }
  if (sp->deblocking) {
    DVC_CHECK(sp->intra_qp >= 0 && sp->intra_qp <= 51, DVC_EINVAL, "WZ frame: intra_qp must be 0..51");
    j->deblock = true;
    if (sp->past_is_wz && sp->past_block_mode) j->pastMode.assign(sp->past_block_mode, sp->past_block_mode + n8);
  }
  if (sp->block_mode || sp->mode_bits) {
    DVC_CHECK(sp->mode_bits >= 2, DVC_EINVAL, "WZ frame: mode_bits counts at least the two mode flags");
    j->bitrate0 += sp->mode_bits - 2;  // band_plan counted one bit per mode flag
  }
  for (auto &b : j->bands) j->nbp += b.qbits;
  j->crc.assign(sp->crc, sp->crc + j->nbp);
  if (!e->deviceBuffers) j->acc.assign(sp->acc, sp->acc + (size_t)j->nbp * (j->N / 8));
  j->bytes = job_bytes(*j);
  j->pastKey = pastKey ? pastKey : (uint64_t)(uintptr_t)sp->past;
  j->futKey = futKey ? futKey : (uint64_t)(uintptr_t)sp->fut;
  j->ticket = e->nextTicket++;
  e->pending.push_back(j);
  if (ticket) *ticket = j->ticket;
  return DVC_OK;
}

static void engine_destroy(dvc_engine *e);

static int engine_create(dvc_ctx *c, const dvc_engine_params *p, dvc_engine **out) {
  DVC_CHECK(c && out, DVC_EINVAL, "dvc_engine_create: NULL argument");
  DVC_HIP(hipSetDevice(c->device));
  int rc = ensure_cost_tables(c);
  if (rc) return rc;
  dvc_engine *e = new dvc_engine();
  e->c = c;
  e->maxInflight = (p && p->max_inflight > 0) ? p->max_inflight : 64;
  e->sequentialBp = p && p->sequential_bp;
  e->graphs = !(p && p->eager_steps);
  e->deviceBuffers = p && p->device_buffers;
  e->laneSteps = p ? p->lane_steps : 0;
  size_t freeB = 0, totalB = 0;
  (void)hipMemGetInfo(&freeB, &totalB);
  e->budget = (p && p->hbm_budget_bytes) ? (size_t)p->hbm_budget_bytes : (size_t)(0.8 * (double)freeB);
  // lanes: as requested, else one per 16 frames in flight, at most 4
  int nl = (p && p->lanes > 0) ? p->lanes : std::min(4, std::max(1, e->maxInflight / 16));
  nl = std::min(nl, e->maxInflight);
  e->lanes.resize(nl);
  // admission staging: frame initialisations, then up to two key frames per frame
  e->admitKeysOff = align_up(sizeof(JobInitDev) * (size_t)e->maxInflight);
  e->admitBytes = e->admitKeysOff + align_up(sizeof(KeyFrameJob) * 2 * (size_t)e->maxInflight);
  hipError_t err = hipHostMalloc((void **)&e->h_admit, e->admitBytes, hipHostMallocDefault);
  if (err == hipSuccess) err = hipMalloc((void **)&e->d_admit, e->admitBytes);
  for (int k = 0; k < nl && err == hipSuccess; k++) {
    Lane &l = e->lanes[k];
    l.cap = (e->maxInflight + nl - 1) / nl;
    l.groupCap = (c->bp_mode == 1 ? 2 * kMaxCodes : kMaxCodes) * l.cap + 1;  // ParallelMode 1 keeps two decision sets per code
    const size_t M = l.cap;
    const size_t oFrames = 0, oLists = align_up(oFrames + sizeof(FrameDev) * M), oLinks = align_up(oLists + sizeof(int) * 5 * M),
                 oPlanes = align_up(oLinks + sizeof(BpFrameLink) * M), oSegs = align_up(oPlanes + sizeof(PlaneDesc) * M),
                 total = align_up(oSegs + sizeof(CopySeg) * 4 * M);
    err = hipStreamCreateWithFlags(&l.s, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipHostMalloc((void **)&l.h_stage, total, hipHostMallocDefault);
    if (err == hipSuccess) err = hipMalloc((void **)&l.d_stage, total);
    if (err == hipSuccess) {
      l.stageBytes = total;
      memset(l.h_stage, 0, total);
      l.h_frames = (FrameDev *)(l.h_stage + oFrames); l.d_frames = (FrameDev *)(l.d_stage + oFrames);
      l.h_lists = (int *)(l.h_stage + oLists); l.d_lists = (int *)(l.d_stage + oLists);
      l.h_links = (BpFrameLink *)(l.h_stage + oLinks); l.d_links = (BpFrameLink *)(l.d_stage + oLinks);
      l.h_planes = (PlaneDesc *)(l.h_stage + oPlanes); l.d_planes = (PlaneDesc *)(l.d_stage + oPlanes);
      l.h_segs = (CopySeg *)(l.h_stage + oSegs); l.d_segs = (CopySeg *)(l.d_stage + oSegs);
    }
    if (err == hipSuccess) err = hipMalloc((void **)&l.d_slots, (size_t)2 * l.groupCap * 6336);
    if (err == hipSuccess) err = hipMalloc((void **)&l.d_counters, 6 * sizeof(int));
  }
  if (err != hipSuccess) {
    set_error(std::string("dvc_engine_create: ") + hipGetErrorString(err));
    engine_destroy(e);
    return DVC_ENOMEM;
  }
  *out = e;
  return DVC_OK;
}

static void engine_destroy(dvc_engine *e) {
  if (!e) return;
  (void)hipSetDevice(e->c->device);
  (void)hipStreamSynchronize(e->c->stream);
  for (auto &l : e->lanes)
    if (l.s) (void)hipStreamSynchronize(l.s);
  for (Job *j : e->pending) delete j;
  for (auto &l : e->lanes) {
    for (Job *j : l.inflight) { if (j->mem) (void)hipFree(j->mem); delete j; }
    for (Job *j : l.finished) { if (j->mem) (void)hipFree(j->mem); delete j; }
    for (auto &g : l.graphs) (void)hipGraphExecDestroy(g.second);
    void *ptrs[] = {l.d_stage, l.d_out, l.d_slots, l.d_counters};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
    if (l.h_stage) (void)hipHostFree(l.h_stage);
    if (l.h_out) (void)hipHostFree(l.h_out);
    if (l.s) (void)hipStreamDestroy(l.s);
  }
  for (auto &kv : e->keys) (void)hipFree(kv.second.k.full);
  for (auto &kv : e->pool) (void)hipFree(kv.second);
  if (e->d_admit) (void)hipFree(e->d_admit);
  if (e->h_admit) (void)hipHostFree(e->h_admit);
  delete e;
}

}  // namespace dvc

extern "C" {

int dvc_engine_create(dvc_ctx *c, const dvc_engine_params *p, dvc_engine **out) { return engine_create(c, p, out); }
int dvc_engine_destroy(dvc_engine *e) { engine_destroy(e); return DVC_OK; }
int dvc_engine_submit(dvc_engine *e, const dvc_wz_frame *f, uint64_t past_key, uint64_t fut_key, uint64_t *ticket) {
  DVC_CHECK(e, DVC_EINVAL, "dvc_engine_submit: NULL engine");
  return engine_submit(e, f, past_key, fut_key, ticket);
}
int dvc_engine_step(dvc_engine *e, int *completed) {
  DVC_CHECK(e && completed, DVC_EINVAL, "dvc_engine_step: NULL argument");
  DVC_HIP(hipSetDevice(e->c->device));
  return engine_step(e, completed);
}
int dvc_engine_status(dvc_engine *e, int *queued, int *inflight, uint64_t *hbm_used) {
  DVC_CHECK(e, DVC_EINVAL, "dvc_engine_status: NULL engine");
  if (queued) *queued = (int)e->pending.size();
  if (inflight) *inflight = e->inflight();
  if (hbm_used) *hbm_used = e->used;
  return DVC_OK;
}
int dvc_engine_done(dvc_engine *e, uint64_t ticket, int *done) {
  DVC_CHECK(e && done, DVC_EINVAL, "dvc_engine_done: NULL argument");
  auto it = e->doneTickets.find(ticket);
  *done = it != e->doneTickets.end();
  if (*done) e->doneTickets.erase(it);
  return DVC_OK;
}
int dvc_engine_release_key(dvc_engine *e, uint64_t key) {
  DVC_CHECK(e, DVC_EINVAL, "dvc_engine_release_key: NULL engine");
  auto it = e->keys.find(key);
  if (it != e->keys.end()) it->second.named = false;
  trim_keys(e, 0);
  return DVC_OK;
}

int dvc_side_info_batch(dvc_ctx *c, int n, const dvc_si_request *reqs) {
  DVC_CHECK(c, DVC_EINVAL, "dvc_side_info_batch: NULL context");
  DVC_CHECK(n >= 0 && (n == 0 || reqs), DVC_EINVAL, "dvc_side_info_batch: bad request list");
  for (int i = 0; i < n; i++) {
    const dvc_si_request &r = reqs[i];
    DVC_CHECK((r.width == 176 && r.height == 144) || (r.width == 352 && r.height == 288), DVC_EINVAL,
              "dvc_side_info_batch: size must be QCIF or CIF");
    DVC_CHECK(r.qindex >= 1 && r.qindex <= 8, DVC_EINVAL, "dvc_side_info_batch: qindex must be 1..8");
    DVC_CHECK(r.past && r.fut && r.si, DVC_EINVAL, "dvc_side_info_batch: NULL pointer");
    DVC_CHECK(r.past_idx < r.cur_idx && r.cur_idx < r.fut_idx, DVC_EINVAL,
              "dvc_side_info_batch: need past_idx < cur_idx < fut_idx");
  }
  if (n == 0) return DVC_OK;
  DVC_HIP(hipSetDevice(c->device));
  int rc = ensure_cost_tables(c);
  if (rc) return rc;
  hipStream_t s = c->stream;
  // one allocation: per request the two key frames (full, low-pass, 2x), SI,
  // residue, its transform, alpha and the motion-field scratch
  std::vector<size_t> base(n + 1, 0);
  int maxN = 0, maxMB = 0;
  for (int i = 0; i < n; i++) {
    const size_t FS = (size_t)reqs[i].width * reqs[i].height;
    size_t t = 0;
    for (size_t b : {6 * FS, 6 * FS, FS, 4 * FS, 4 * FS, 4 * FS, sizeof(MVd) * (FS / 256), sizeof(MVd) * (FS / 64),
                     sizeof(MVd) * (FS / 64), sizeof(MVd) * (FS / 256), 16 * (FS / 256), 8 * (FS / 64)})
      t = align_up(t + b);
    base[i + 1] = base[i] + t;
    maxN = std::max(maxN, (int)(FS / 16));
    maxMB = std::max(maxMB, (int)(FS / 256));
  }
  // the context keeps the allocation for later calls (grown when too small)
  const size_t need = base[n] + sizeof(FrameDev) * n + sizeof(int) * n;
  if (c->si_scratch_bytes < need) {
    if (c->si_scratch) DVC_HIP(hipFree(c->si_scratch));
    c->si_scratch = nullptr;
    c->si_scratch_bytes = 0;
    DVC_HIP(hipMalloc(&c->si_scratch, need));
    c->si_scratch_bytes = need;
  }
  uint8_t *mem = (uint8_t *)c->si_scratch;
  FrameDev *d_frames = (FrameDev *)(mem + base[n]);
  int *d_list = (int *)(d_frames + n);
  std::vector<FrameDev> hf(n);
  std::vector<int> list(n);
  auto run = [&]() -> int {
    for (int i = 0; i < n; i++) {
      const dvc_si_request &r = reqs[i];
      const int W = r.width, H = r.height;
      const size_t FS = (size_t)W * H;
      uint8_t *p = mem + base[i];
      size_t off = 0;
      auto take = [&](size_t bytes) { uint8_t *q = p + off; off = align_up(off + bytes); return q; };
      uint8_t *kp = take(6 * FS), *kf = take(6 * FS);
      FrameDev &f = hf[i];
      f = FrameDev{};
      f.W = W; f.H = H; f.N = (int)(FS / 16); f.qindex = r.qindex;
      f.iPC = (double)(r.cur_idx - r.past_idx) / (double)(r.fut_idx - r.past_idx);
      f.iFC = (double)(r.fut_idx - r.cur_idx) / (double)(r.fut_idx - r.past_idx);
      f.past = kp; f.past_lp = kp + FS; f.past_half = kp + 2 * FS;
      f.fut = kf; f.fut_lp = kf + FS; f.fut_half = kf + 2 * FS;
      f.si = take(FS);
      f.residue = (int *)take(4 * FS);
      f.sideTrans = (int *)take(4 * FS);
      f.alpha = (float *)take(4 * FS);
      f.mbmv = (MVd *)take(sizeof(MVd) * (FS / 256));
      f.bmv = (MVd *)take(sizeof(MVd) * (FS / 64));
      f.bmv_tmp = (MVd *)take(sizeof(MVd) * (FS / 64));
      f.fwd = (MVd *)take(sizeof(MVd) * (FS / 256));
      f.holes = (double *)take(16 * (FS / 256));
      f.mse = (double *)take(8 * (FS / 64));
      list[i] = i;
      DVC_HIP(hipMemcpyAsync(kp, r.past, FS, hipMemcpyHostToDevice, s));
      DVC_HIP(hipMemcpyAsync(kf, r.fut, FS, hipMemcpyHostToDevice, s));
      DVC_HIP(hipMemsetAsync(f.alpha, 0, 4 * FS, s));
      TIMED_LAUNCH(c, "keyframe_filters", s, launch_keyframe_filters(kp, kp + FS, kp + 2 * FS, W, H, s));
      TIMED_LAUNCH(c, "keyframe_filters", s, launch_keyframe_filters(kf, kf + FS, kf + 2 * FS, W, H, s));
    }
    DVC_HIP(hipMemcpyAsync(d_frames, hf.data(), sizeof(FrameDev) * n, hipMemcpyHostToDevice, s));
    DVC_HIP(hipMemcpyAsync(d_list, list.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    CostTables ct{c->d_costFme, c->d_costUsi};
    TIMED_LAUNCH(c, "create_si", s, launch_create_si(d_frames, d_list, n, maxMB, ct, s));
    TIMED_LAUNCH(c, "noise_param", s, launch_noise_param(d_frames, d_list, n, maxN, s));
    for (int i = 0; i < n; i++) {
      const dvc_si_request &r = reqs[i];
      const size_t FS = (size_t)r.width * r.height;
      DVC_HIP(hipMemcpyAsync(r.si, hf[i].si, FS, hipMemcpyDeviceToHost, s));
      if (r.residue) DVC_HIP(hipMemcpyAsync(r.residue, hf[i].residue, 4 * FS, hipMemcpyDeviceToHost, s));
      if (r.alpha) DVC_HIP(hipMemcpyAsync(r.alpha, hf[i].alpha, 4 * FS, hipMemcpyDeviceToHost, s));
    }
    DVC_HIP(hipStreamSynchronize(s));
    return DVC_OK;
  };
  rc = run();
  (void)hipStreamSynchronize(s);
  return rc;
}

int dvc_decode_wz_frames(dvc_ctx *c, int n, const dvc_wz_frame *frames, int batched) {
  DVC_CHECK(c, DVC_EINVAL, "dvc_decode_wz_frames: NULL context");
  DVC_CHECK(n >= 0 && (n == 0 || frames), DVC_EINVAL, "dvc_decode_wz_frames: bad frame list");
  if (n == 0) return DVC_OK;
  dvc_engine_params p{};
  p.max_inflight = batched ? n : 1;
  p.sequential_bp = batched ? 0 : 1;
  p.lane_steps = 1 << 20;  // every frame is submitted up front: lanes never wait for each other
  dvc_engine *e = nullptr;
  int rc = engine_create(c, &p, &e);
  if (rc) return rc;
  for (int i = 0; i < n && !rc; i++) rc = engine_submit(e, &frames[i], 0, 0, nullptr);
  int left = n;
  while (!rc && left > 0) {
    int done = 0;
    rc = engine_step(e, &done);
    left -= done;
    if (!rc && done == 0 && e->inflight() == 0 && !e->pending.empty()) {
      set_error("dvc_decode_wz_frames: frame does not fit the HBM budget");
      rc = DVC_ENOMEM;
    }
  }
  engine_destroy(e);
  return rc;
}

}  // extern "C"
