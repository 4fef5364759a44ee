// engine_kernels.h -- batched set-up and tear-down work of the serving engine
// (engine_kernels.hip): one launch initialises every frame admitted in a
// step, filters every new key frame, or gathers every finished frame's
// results, instead of a copy or fill per buffer per frame.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "ldpca_kernels.h"

namespace dvc {

// Device state of one admitted frame to initialise.
struct JobInitDev {
  int *reconsTrans;           // fs ints set to 0 (no bit decoded yet)
  int fs;
  uint8_t *carry;             // N bytes set to 0xAA (ParallelMode 1's carried decisions)
  int N;
  PlaneState *state;          // nbp rate-search states: start from the entropy, nothing accepted
  int4 *rec;                  // kMaxCodes * nbp run records: not run
  int nbp;
  int numCodes;               // codes of the frame's ladder ("nothing accepted" = numCodes - 1)
  const uint8_t *accSrc;      // accumulated syndromes to copy (device), or null
  uint8_t *accDst;
  int accBytes;
};

// One key frame to bring in: the image is copied from src to full unless
// src is full, then the 3x3 low-pass and the 2x interpolated image are
// computed from it.
struct KeyFrameJob {
  const uint8_t *src;
  uint8_t *full, *lp, *half;
  int W, H;
};

// A device-to-device byte copy.
struct CopySeg {
  const uint8_t *src;
  uint8_t *dst;
  size_t bytes;
};

hipError_t launch_job_init(const JobInitDev *items, int n, hipStream_t s);
hipError_t launch_keyframe_batch(const KeyFrameJob *jobs, int n, int maxW, int maxH, hipStream_t s);
hipError_t launch_copy_segments(const CopySeg *segs, int n, hipStream_t s);

}  // namespace dvc
