// frame_kernels.h -- per-WZ-frame device state and the launchers of the side
// information, transform, noise-model, motion-learning and reconstruction
// kernels (frame_kernels.hip).  Every kernel takes the device array of frame
// descriptors and a list of the frames it works on, so one launch serves a
// batch of frames of any mix of sizes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dvc {

struct MVd {                 // MV of global.h:170-174: past, future, vector (x, y)
  double px, py, fx, fy, vx, vy;
};

struct FrameDev {
  // geometry
  int W, H, N;               // N = number of 4x4 blocks = LDPCA code length
  int qindex;
  // block modes per 8x8 block (BlockModeFlag: 0 WZ, 1 intra, 2 skip), or
  // nullptr when every block is a WZ block; only WZ blocks take decoded bits
  // and are reconstructed, refined and modelled
  const uint8_t *mode;
  // intra blocks: the host-parsed dequantised coefficients, 16 per 4x4 block
  // (row-major, zero outside intra blocks), or nullptr without intra blocks
  const int *intraCoeffs;
  // deblocking (DeblockingMode): per coefficient position whether
  // reconstruction changed it (deblockFlag), per edge the filter mode and
  // strength, the past reference's block modes when it is a WZ frame, the
  // filter's QP and 2^(QP/6), and the coefficient weights Qtable[i] / sum
  int deblock, pastIsWz, intraQp;
  double pow2qp;
  double dbWeight[16];
  uint8_t *dbFlag;
  uint8_t *dbEdgeMode;
  double *dbEdgeBs;
  const uint8_t *pastMode;
  double iPC, iFC;           // (cur-past)/(fut-past), (fut-cur)/(fut-past)
  // key frames: full resolution, 3x3 low-pass, 2x interpolated (2W x 2H)
  const uint8_t *past, *fut, *past_lp, *fut_lp, *past_half, *fut_half;
  // working state
  uint8_t *si;               // W*H side information
  int *residue;              // W*H (fut - past) >> 1 of the MC step
  int *sideTrans;            // W*H SI coefficients (residue transform first, then ML-refined SI)
  int *reconsTrans;          // W*H decoded quantisation indices
  float *alpha;              // W*H Laplacian parameter per coefficient
  int *oct;                  // (H-3)*(W-3)*16 over-complete transform of the SI
  int *octBand;              // (H-3)*(W-3) one coefficient of oct per position: the index of the band
                             // whose noise distribution comes next
  float *smf;                // N*81 statistical motion field probabilities
  float *dist;               // N*distStride noise distribution of the current band
  int distStride;            // range (DC) or 2*range+1 (AC)
  float *LR;                 // N likelihood ratios of the current bitplane
  uint8_t *decoded;          // N decoded bits of the current bitplane
  uint8_t *recons;           // W*H reconstruction
  MVd *mbmv;                 // macroblock MVs (W/16*H/16)
  MVd *bmv;                  // 8x8 block MVs (W/8*H/8)
  MVd *bmv_tmp;              // 8x8 MVs after the vector-median filter
  MVd *fwd;                  // forward-ME result per macroblock
  double *holes;             // 2 per macroblock
  double *mse;               // per 8x8 block
  // per-band parameters (set by the host before the band's kernels)
  int band;                  // zig-zag band number
  int index;                 // coefficient index in the 4x4 block (zigzag[band])
  int range, step, levelShift, bit;
};

// Cost-function factors 1 + K*sqrt(x^2 + y^2) (sideInfoCreation.cpp:10-12)
// evaluated on the host for the integer offsets of forwardME (|x|,|y| <= 32) and
// updateSI (x, y even, |x|,|y| <= 64), 65x65 each.
struct CostTables {
  const double *fme;   // [(dy+32)*65 + dx+32]
  const double *usi;   // [(dy+32)*65 + dx+32] for mv = 2*(dx, dy)
};

// Intra blocks: DC prediction + inverse transform of the parsed coefficients
// into recons, then the intra 8x8 blocks into the side information (one
// workgroup per frame for the first; frames without intra blocks return).
hipError_t launch_intra(const FrameDev *F, const int *list, int n, int maxBlk88, hipStream_t s);
// The deblocking filter of the frames whose `deblock` is set, in place on recons.
hipError_t launch_deblock(const FrameDev *F, const int *list, int n, int maxW, int maxH, hipStream_t s);
// A WZ frame's skip blocks: the past key frame's pixels (copySkipBlkToRecons,
// skipBlock.cpp:3-22), for the frames of the list that have a mode map.
hipError_t launch_skip_copy(const FrameDev *F, const int *list, int n, int maxBlk88, hipStream_t s);
hipError_t launch_keyframe_filters(const uint8_t *in, uint8_t *lp, uint8_t *half, int W, int H, hipStream_t s);
hipError_t launch_create_si(const FrameDev *F, const int *list, int n, int maxMB, CostTables ct, hipStream_t s);
// forwardME alone (sideInfoCreation.cpp:61-145): macroblock MVs (mbmv, unscaled) from past_lp/fut_lp.
hipError_t launch_forward_me(const FrameDev *F, const int *list, int n, int maxMB, CostTables ct, hipStream_t s);
// Writes octBand from oct for coefficient `index` (the per-band entry points,
// whose over-complete transform comes from the caller).
hipError_t launch_oct_plane(const FrameDev *F, const int *list, int n, int maxPos, int index, hipStream_t s);
hipError_t launch_noise_param(const FrameDev *F, const int *list, int n, int maxN, hipStream_t s);
hipError_t launch_oct(const FrameDev *F, const int *list, int n, int maxPos, int initSideTrans, hipStream_t s);
hipError_t launch_noise_dist(const FrameDev *F, const int *list, int n, int maxN, int maxStride, hipStream_t s);
hipError_t launch_condbit(const FrameDev *F, const int *list, int n, int maxN, hipStream_t s);
struct BpFrameLink;
hipError_t launch_put_bitplane(const FrameDev *F, const int *list, const BpFrameLink *links, int n, int maxN,
                               hipStream_t s);
hipError_t launch_reconstruct(const FrameDev *F, const int *list, int n, int maxN, hipStream_t s);
hipError_t launch_inverse_transform(const FrameDev *F, const int *list, int n, int maxN, hipStream_t s);
hipError_t launch_update_si(const FrameDev *F, const int *list, int n, int maxN, CostTables ct, hipStream_t s);
hipError_t launch_smf_init(const FrameDev *F, const int *list, int n, int maxN, hipStream_t s);
hipError_t launch_update_smf(const FrameDev *F, const int *list, int n, int maxN, hipStream_t s);
hipError_t launch_update_si_ml(const FrameDev *F, const int *list, int n, int maxBw, int maxBh, hipStream_t s);

}  // namespace dvc
