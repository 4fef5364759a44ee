// dvc_internal.h -- shared host-side state of the gfx950 decoding library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <mutex>
#include <string>
#include <vector>
#include "../../include/dvc_accel.h"

namespace dvc {
struct ColEdges;  // ldpca_kernels.h
}

namespace dvc {

void set_error(const std::string &msg);

#define DVC_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      dvc::set_error(std::string(#call) + ": " + hipGetErrorString(e_));               \
      return DVC_EDEVICE;                                                              \
    }                                                                                  \
  } while (0)

#define DVC_CHECK(cond, code, msg)                                                     \
  do {                                                                                 \
    if (!(cond)) { dvc::set_error(msg); return (code); }                               \
  } while (0)

// The BP kernels' threads per workgroup on an irregular ladder (its
// column-to-thread table is built for this count).
constexpr int kIrregThreads = 256;

// Host copy of one ladder (all codes of one code length) and its device image.
struct Ladder {
  int N = 0, E = 0, period = 0, numCodes = 0;
  bool regular = true;             // every column has degree 3 (the Reg*.lad ladders)
  std::vector<int> jc;             // N+1 column pointers
  std::vector<int> encIr;          // E encoder rows
  std::vector<int> bitsPerPeriod;  // numCodes
  std::vector<int> idx;            // numCodes*period
  std::vector<int> syndLength;     // numCodes
  std::vector<float> codeRate;     // numCodes
  std::vector<int> decIr;          // numCodes*E decoder rows per edge
  std::vector<uint8_t> invH;       // N * invRowBytes, MSB-first
  int invRowBytes = 0;             // 4*ceil(N/32)
  std::vector<uint8_t> crcBasis;   // N: CRC-8 of the message with only bit k set

  // Device image, per BP code c (0..numCodes-2):
  //   d_colRows[c*N + k]       column k's edges 3k..3k+2, one word each:
  //                            row | slot << 15, slot = the edge's message
  //                            position in row order (rowPtr[row] + its rank
  //                            among the row's edges in ascending edge order)
  //   d_rowPtr[c*(N+1) + r]    number of edges in rows below r (r <= m): row
  //                            r's messages sit at rowPtr[r] + rank
  //   d_tx[c*N + k]            accumulated-syndrome position of syndrome bit k
  ColEdges *d_colRows = nullptr;
  // Irregular ladders (column degrees differ) instead of d_colRows:
  //   d_colEdges[c*E + l]   edge l (column order) as row | rank << 16
  //   d_jc[k]               column pointers (N+1)
  //   d_colSlots[t*S + i]   the i-th column thread t of a kIrregThreads
  //                         workgroup updates (-1: none), S = colSlotsPerThread,
  //                         columns dealt to threads so each thread's edge
  //                         count is nearly equal
  uint32_t *d_colEdges = nullptr;
  int *d_jc = nullptr;
  int16_t *d_colSlots = nullptr;
  int colSlotsPerThread = 0;
  uint16_t *d_rowPtr = nullptr, *d_tx = nullptr;
  int *d_syndLength = nullptr;
  float *d_codeRate = nullptr;
  uint8_t *d_invH = nullptr, *d_crcBasis = nullptr;
};

// Loads dir/<kind><N>.ladder (kind "reg" or "irreg") and its cached inverse.
int load_ladder(const std::string &dir, const char *kind, int N, Ladder &L);
// BP launch shape of the context's tuning: threads per workgroup and the
// persistent grid for `items` queued (bitplane, code) items.
int bp_threads(const dvc_ctx *c, int N);
int bp_grid(const dvc_ctx *c, int N, int items);
int upload_ladder(Ladder &L);
// Quantisation table entry Qtable[q][index] (reconstruction.cpp:9-78), 0 outside 1..8 x 0..15.
int qtable_entry(int q, int index);
// QuantizationTableBits[q][index] (reconstruction.cpp:44-53) and zigzag[band] (global.cpp:36-41).
int qbits_entry(int q, int index);
int zigzag_entry(int band);
// Motion-cost factor tables of the context (made on first use).
int ensure_cost_tables(dvc_ctx *c);
// Per-band entry points' device state (band.hip), freed with the context.
struct BandState;
void band_free(dvc_ctx *c);
// Looks up the allocation-free entry point's scratch array on the context's
// (current) device, once per device (noalloc.hip).
int noalloc_attach(dvc_ctx *c);
void free_ladder(Ladder &L);

}  // namespace dvc

struct dvc_ctx;
// Per-kernel timing with HIP events recorded on the launching stream.
struct KernelTimer {
  std::string name;
  double total_ms = 0.0;
  long launches = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

struct dvc_ctx {
  bool profiling = false;
  std::vector<KernelTimer> timers;
  std::mutex timer_mu;  // engine lanes launch from several host threads
  // Records start/stop events around a launch when profiling is on.
  hipEvent_t timer_begin(const char *name, hipStream_t s, int *slot);
  void timer_end(int slot, hipEvent_t start, hipStream_t s);
  void timer_collect();

  int device = 0;
  int sm_count = 256;
  double *d_costFme = nullptr, *d_costUsi = nullptr;  // motion-cost factor tables
  int max_iter = 50;
  // BP launch configuration (dvc_autotune / dvc_set_tuning): threads per
  // workgroup for each code length, QCIF workgroups per CU, and the interval
  // in iterations at which a running speculative item polls for cancellation
  // (defaults: dvc_autotune's choices on MI355X)
  int bp_threads_cif = 1024;
  int bp_threads_qcif = 512;
  int bp_wg_per_cu_qcif = 4;
  int bp_poll = 1;
  int bp_mode = 0;  // 0: ParallelMode 0 BP semantics, 1: ParallelMode 1 (early jump on repeated decisions)
  int rate_step = 0;  // 1: DYNAMICSTEP rate search (dvc_params.rate_step)
  dvc::Ladder lad1584, lad6336;
  hipStream_t stream = nullptr;
  void *si_scratch = nullptr;  // dvc_side_info_batch's device buffers, kept between calls
  size_t si_scratch_bytes = 0;
  void *ld_dev = nullptr;      // the LDPCA entry points' device workspace, kept between calls
  size_t ld_dev_bytes = 0;
  void *ld_host = nullptr;     // its pinned host staging (inputs up, results back)
  size_t ld_host_bytes = 0;
  void *noalloc_scratch = nullptr;  // device address of noalloc.hip's scratch array on this device
  dvc::BandState *band = nullptr;   // per-band entry points' frame state (dvc_band_frame_init)
  dvc::Ladder *ladder(int N) { return N == 1584 ? &lad1584 : (N == 6336 ? &lad6336 : nullptr); }
};
