// ldpca_kernels.h -- argument blocks shared by the LDPCA kernels and their
// host-side launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dvc {

struct Ladder;

// Per-bitplane mutable state on the device.
struct PlaneState {
  int start;      // nCodeStart (-1: computed by the CE kernel)
  int best;       // lowest code that converged with a matching CRC (numCodes-1 = none)
  float entropy;  // H(X|Y) from ComputeCE
  int run;        // first code of the current launch's queue (start, or the code a
                  // ParallelMode 1 resolution resumes at; numCodes = nothing left)
  int taint;      // 1: some LLR of the bitplane is NaN (set by the CE kernel): a likelihood
                  // ratio condBitProb formed as 0/0 when both of a block's interval sums
                  // underflow to zero (noiseModel.cpp:295-303)
};

struct PlaneDesc {
  int N;
  const float *LR;     // N likelihood ratios
  float *LLR;          // N log-likelihood ratios (written by the CE kernel)
  const uint8_t *acc;  // N/8 accumulated-syndrome bytes
  uint32_t crc;        // encoder CRC-8
  PlaneState *state;
  int4 *rec;           // [numCodes] per code: iterations, status, decision slot, used_inverse
  // ParallelMode 1 only (NULL otherwise):
  uint8_t *carry;      // N decision bytes carried into code `run` (any byte other than 0/1 never matches)
  int *bcase;          // [numCodes] decision slot + 1 of a code's iteration-3 decisions when they
                       // repeat its first three (see bp_persistent_kernel), else 0
};

// What k_put_bitplane needs to find the accepted decisions of a bitplane.
struct BpFrameLink {
  const PlaneState *state;
  const int4 *records;
  const uint8_t *slots;
  int slot_capacity;
  uint8_t *carry;  // ParallelMode 1: receives the accepted decisions (the next bitplane's carry)
};

// Run records per bitplane in the frame decoder's buffers: the most codes a
// ladder has (the irregular QCIF ladder's 66; the regular ladders have 65).
constexpr int kMaxCodes = 66;

// A degree-3 column's edges for the BP kernel: row | message slot << 15 each.
struct ColEdges {
  uint32_t e[3];
};

struct LadderDev {
  int N, E, numCodes;
  int regular;           // every column has degree 3 (colRows), else colEdges/jc/colSlots
  const ColEdges *colRows;  // per code and column: its three edges as row | message slot << 15
  const uint32_t *colEdges;  // irregular: per code and edge (column order) row | rank << 16
  const int *jc;             // irregular: column pointers
  const int16_t *colSlots;   // irregular: columns of each of kIrregThreads threads (-1: none)
  int colSlotsPerThread;
  const uint16_t *rowPtr, *tx;
  const int *syndLength;
  const uint8_t *invH, *crcBasis;
};

struct BpArgs {
  LadderDev lad;
  const PlaneDesc *planes;
  int nplanes;
  int *next;          // queue head (reset to 0 before a launch)
  int total;          // nplanes * max_offset
  int max_offset;     // codes above start tried per bitplane
  int max_iter;
  uint8_t *slots;     // accepted decisions, N bytes per slot
  int *slot_count;
  int slot_capacity;
  int poll;           // iterations between polls of the bitplane's best code (>= 1)
  int mode;           // 0: beliefPropagation, 1: beliefPropagation_OpenMP (ParallelMode 1)
  int rate_step;      // 1: only the codes the DYNAMICSTEP search tries (dynstep_tries)
};

// DYNAMICSTEP rate search (LdpcaDecoder.cpp:19-21, 669-672, 760-768): after
// the r-th failed request (r = 0, 1, ...) the search moves 1 + {0,1,2,3}[min(r,3)]
// codes up, so it tries the codes start + 0, 1, 3, 6, 10, 14, 18, ... below
// the full rate, then the full rate.
__host__ __device__ inline bool dynstep_tries(int off) {
  return off == 0 || off == 1 || off == 3 || off == 6 || (off >= 10 && (off - 10) % 4 == 0);
}
// The offset from the start code tried after offset `off` fails.
__host__ __device__ inline int dynstep_next(int off) {
  return off < 1 ? 1 : off < 3 ? 3 : off < 6 ? 6 : off < 10 ? 10 : off + 4;
}

// record.y status values
enum { REC_FAILED = 0, REC_CRC_MISMATCH = 1, REC_ACCEPTED = 2, REC_CANCELLED = 3, REC_NOT_RUN = -1 };

LadderDev ladder_dev(const Ladder &L);
hipError_t launch_bp_any(int N, int threads, const BpArgs &a, int grid, hipStream_t s);
hipError_t launch_zero_ints(int *p, int n, hipStream_t s);
hipError_t launch_mode1_resolve(const PlaneDesc *planes, int n, int numCodes, int N, const uint8_t *slots,
                                int slot_capacity, int *rerun_count, hipStream_t s);
hipError_t launch_ce_llr(const PlaneDesc *d_planes, int n, int fixed_start, const Ladder &l1584,
                         const Ladder &l6336, hipStream_t s);
// BP on an irregular ladder (bp_irregular_kernel, kIrregThreads per workgroup).
hipError_t launch_bp_irregular(const BpArgs &a, int grid, hipStream_t s);
hipError_t launch_invh(const PlaneDesc *d_planes, int n, const LadderDev &L, uint8_t *slots, int *slot_count,
                       int slot_capacity, hipStream_t s);
// Copies bitplane p's accepted decisions (N bytes, N a multiple of 16) to
// out + p * N for every p < n; one with no accepted record in range is skipped.
hipError_t launch_gather_accepted(const PlaneDesc *d_planes, int n, int N, const uint8_t *slots, int slot_capacity,
                                  uint8_t *out, hipStream_t s);

}  // namespace dvc
