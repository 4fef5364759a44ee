// intra_deblock.hip -- intra blocks and the deblocking filter of a WZ frame
// on the device.
//
// Intra blocks (intraRecon, intra.cpp:240-258): 4x4 DC prediction from the
// left column and upper row of neighbouring intra blocks, the H.264 integer
// inverse transform of the dequantised coefficients the host parsed
// (intra_dec.cpp), clip.  A block reads only blocks to its left and above,
// so one workgroup per frame sweeps the anti-diagonals of the 4x4 block grid
// (all blocks of a diagonal at once, a barrier between diagonals), which
// keeps the reference's raster-order results.  The intra 8x8 blocks then
// replace the side information (copyIntraBlkToSideInfo, intra.cpp:260-288).
//
// Deblocking (deblock, deblock.cpp:346-386): every edge's filter mode and
// strength are decided first from the unfiltered reconstruction
// (deblock_44_decision, :208-344); then the vertical edges are filtered, one
// thread per pixel row walking its edges left to right, and then the
// horizontal edges, one thread per pixel column walking down -- the order in
// which the reference's raster loops touch each row and column.
#include "frame_kernels.h"

namespace dvc {

constexpr uint8_t kModeWz = 0, kModeIntra = 1, kModeSkip = 2;

__device__ __forceinline__ uint8_t mode88(const FrameDev &f, int bx, int by) {  // by 4x4 block coordinates
  return f.mode ? f.mode[(by >> 1) * (f.W >> 3) + (bx >> 1)] : kModeWz;
}

// inverse44_intra (intra.cpp:36-78)
__device__ void inverse44_intra(int *b) {
  int t[16];
  for (int c = 0; c < 4; c++) {
    const int x0 = b[c], x1 = b[4 + c], x2 = b[8 + c], x3 = b[12 + c];
    t[c] = x0 + x1 + x2 + (x3 >> 1);
    t[4 + c] = x0 + (x1 >> 1) - x2 - x3;
    t[8 + c] = x0 - (x1 >> 1) - x2 + x3;
    t[12 + c] = x0 - x1 + x2 - (x3 >> 1);
  }
  for (int r = 0; r < 4; r++) {
    const int y0 = t[4 * r], y1 = t[4 * r + 1], y2 = t[4 * r + 2], y3 = t[4 * r + 3];
    b[4 * r] = (y0 + y1 + y2 + (y3 >> 1)) >> 6;
    b[4 * r + 1] = (y0 + (y1 >> 1) - y2 - y3) >> 6;
    b[4 * r + 2] = (y0 - (y1 >> 1) - y2 + y3) >> 6;
    b[4 * r + 3] = (y0 - y1 + y2 - (y3 >> 1)) >> 6;
  }
}

// One workgroup per frame of the list (frames without intra blocks return).
__global__ __launch_bounds__(256) void k_intra_recon(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.x]];
  if (!f.intraCoeffs) return;
  const int bw = f.W >> 2, bh = f.H >> 2, W = f.W;
  for (int d = 0; d < bw + bh - 1; d++) {
    for (int y = threadIdx.x; y < bh; y += blockDim.x) {
      const int x = d - y;
      if (x < 0 || x >= bw || mode88(f, x, y) != kModeIntra) continue;
      // intra4x4_dc_pred (intra.cpp:120-178): neighbours count only when intra
      const bool left = x > 0 && mode88(f, x - 1, y) == kModeIntra;
      const bool up = y > 0 && mode88(f, x, y - 1) == kModeIntra;
      const int top = y * 4, lft = x * 4;
      int s0 = 0;
      if (up)
        for (int n = 0; n < 4; n++) s0 += f.recons[(top - 1) * W + lft + n];
      if (left)
        for (int m = 0; m < 4; m++) s0 += f.recons[(top + m) * W + lft - 1];
      if (up && left) s0 = (s0 + 4) >> 3;
      else if (up || left) s0 = (s0 + 2) >> 2;
      else s0 = 128;
      int b[16];
      const int *c = f.intraCoeffs + 16 * ((size_t)y * bw + x);
      for (int k = 0; k < 16; k++) b[k] = c[k];
      inverse44_intra(b);
      for (int m = 0; m < 4; m++)
        for (int n = 0; n < 4; n++) {
          const int v = b[4 * m + n] + (uint8_t)s0;  // the prediction is stored as unsigned char
          f.recons[(top + m) * W + lft + n] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
        }
    }
    __syncthreads();
  }
}

// copyIntraBlkToSideInfo: one thread per 8x8 block.
__global__ void k_intra_si(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.y]];
  if (!f.intraCoeffs) return;
  const int w8 = f.W >> 3, n8 = w8 * (f.H >> 3);
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n8 || f.mode[k] != kModeIntra) return;
  const int top = (k / w8) * 8, left = (k % w8) * 8;
  for (int m = 0; m < 8; m++)
    for (int n = 0; n < 8; n++) f.si[(top + m) * f.W + left + n] = f.recons[(top + m) * f.W + left + n];
}

// ---- deblocking
// H.264's alpha, beta and tc0 tables indexed by QP (deblock.cpp:12-40).
__constant__ uint8_t c_alpha[52] = {0,  0,  0,  0,  0,  0,  0,  0,   0,   0,   0,   0,   0,   0,   0,   0,   4,   4,
                                    5,  6,  7,  8,  9,  10, 12, 13,  15,  17,  20,  22,  25,  28,  32,  36,  40,  45,
                                    50, 56, 63, 71, 80, 90, 101, 113, 127, 144, 162, 182, 203, 226, 255, 255};
__constant__ uint8_t c_beta[52] = {0, 0, 0, 0, 0, 0, 0, 0,  0,  0,  0,  0,  0,  0,  0,  0,  2,  2,  2,  3,  3,  3,  3,  4,  4,  4,
                                   6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13, 14, 14, 15, 15, 16, 16, 17, 17, 18, 18};
__constant__ uint8_t c_tc0[52][5] = {
    {0, 0, 0, 0, 0},   {0, 0, 0, 0, 0},   {0, 0, 0, 0, 0},   {0, 0, 0, 0, 0},    {0, 0, 0, 0, 0},    {0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0},   {0, 0, 0, 0, 0},   {0, 0, 0, 0, 0},   {0, 0, 0, 0, 0},    {0, 0, 0, 0, 0},    {0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0},   {0, 0, 0, 0, 0},   {0, 0, 0, 0, 0},   {0, 0, 0, 0, 0},    {0, 0, 0, 0, 0},    {0, 0, 0, 1, 1},
    {0, 0, 0, 1, 1},   {0, 0, 0, 1, 1},   {0, 0, 0, 1, 1},   {0, 0, 1, 1, 1},    {0, 0, 1, 1, 1},    {0, 1, 1, 1, 1},
    {0, 1, 1, 1, 1},   {0, 1, 1, 1, 1},   {0, 1, 1, 1, 1},   {0, 1, 1, 2, 2},    {0, 1, 1, 2, 2},    {0, 1, 1, 2, 2},
    {0, 1, 1, 2, 2},   {0, 1, 2, 3, 3},   {0, 1, 2, 3, 3},   {0, 2, 2, 3, 3},    {0, 2, 2, 4, 4},    {0, 2, 3, 4, 4},
    {0, 2, 3, 4, 4},   {0, 3, 3, 5, 5},   {0, 3, 4, 6, 6},   {0, 3, 4, 6, 6},    {0, 4, 5, 7, 7},    {0, 4, 5, 8, 8},
    {0, 4, 6, 9, 9},   {0, 5, 7, 10, 10}, {0, 6, 8, 11, 11}, {0, 6, 8, 13, 13},  {0, 7, 10, 14, 14}, {0, 8, 11, 16, 16},
    {0, 9, 12, 18, 18}, {0, 10, 13, 20, 20}, {0, 11, 15, 23, 23}, {0, 13, 17, 25, 25}};

constexpr int kNoFilter = 0, kWzFilter = 1, kIntraSideFilter = 2;  // NO_F, WZ_mode, ISWZ_mode
constexpr double kBsMax = 0.14;

// deblock_44_decision for the edge between 4x4 blocks (x1, y1) and (x2, y2).
__device__ int deblock_decide(const FrameDev &f, int x1, int y1, int x2, int y2, double &bs) {
  const uint8_t m1 = mode88(f, x1, y1), m2 = mode88(f, x2, y2);
  if (m1 == kModeIntra && m2 == kModeIntra) { bs = 3.0; return kIntraSideFilter; }
  if (m1 == kModeIntra || m2 == kModeIntra) {
    bs = (m1 == kModeWz || m2 == kModeWz) ? 4.0 : 3.0;
    return kIntraSideFilter;
  }
  if (m1 == kModeSkip && m2 == kModeSkip) { bs = 0.0; return kNoFilter; }
  if (m1 == kModeSkip || m2 == kModeSkip) {
    if (f.pastIsWz) {
      // the skip block's mode in the past reference (a WZ frame)
      const int sx = m1 == kModeSkip ? x1 : x2, sy = m1 == kModeSkip ? y1 : y2;
      const uint8_t pm = f.pastMode ? f.pastMode[(sy >> 1) * (f.W >> 3) + (sx >> 1)] : kModeWz;
      bs = pm == kModeIntra ? 2.0 : 1.0;
    } else {
      bs = 2.0;
    }
    return kIntraSideFilter;
  }
  // two WZ blocks: weighted count of coefficients the decoding changed, and
  // whether each block is flat
  double phi[2], chi[2];
  const int bx[2] = {x1, x2}, by[2] = {y1, y2};
  for (int s = 0; s < 2; s++) {
    double p = 0.0;
    int sum = 0, sq = 0;
    for (int m = 0; m < 4; m++) {
      const int idx = (by[s] * 4 + m) * f.W + bx[s] * 4;
      for (int n = 0; n < 4; n++) {
        if (f.dbFlag[idx + n]) p += f.dbWeight[4 * m + n];
        const int v = f.recons[idx + n];
        sum += v;
        sq += v * v;
      }
    }
    const double var = (double)sq / 16.0 - ((double)sum / 16.0) * ((double)sum / 16.0);
    phi[s] = p;
    chi[s] = var < 50.0 ? 1.0 : 0.0;
  }
  const double P = 0.9;
  bs = kBsMax * (P * ((phi[0] + phi[1]) / 2.0) + (1.0 - P) * ((chi[0] + chi[1]) / 2.0));
  return bs > 0.0 ? kWzFilter : kNoFilter;
}

// Decisions of every vertical (dir 0) and horizontal (dir 1) edge, from the
// unfiltered reconstruction.  Edge e of dir 0: blocks (j, i) | (j + 1, i) at
// e = i * (bw - 1) + j; of dir 1: (j, i) over (j, i + 1) at e = i * bw + j.
__global__ void k_deblock_decide(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.y]];
  if (!f.deblock) return;
  const int bw = f.W >> 2, bh = f.H >> 2, nv = (bw - 1) * bh, nh = bw * (bh - 1);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nv + nh) return;
  double bs = 0.0;
  int mode;
  if (e < nv) {
    const int i = e / (bw - 1), j = e % (bw - 1);
    mode = deblock_decide(f, j, i, j + 1, i, bs);
  } else {
    const int k = e - nv, i = k / bw, j = k % bw;
    mode = deblock_decide(f, j, i, j, i + 1, bs);
  }
  f.dbEdgeMode[e] = (uint8_t)mode;
  f.dbEdgeBs[e] = bs;
}

__device__ __forceinline__ int clip3(int lo, int hi, int x) { return x < lo ? lo : (x > hi ? hi : x); }

// deblock_filter (deblock.cpp:42-206) on one line of one edge: p[0..3] are
// L3..L0 and p[4..7] R0..R3, `st` the stride between them.
__device__ void deblock_line(uint8_t *img, size_t at, int st, int mode, double bs, int qp, double pow2qp) {
  uint8_t *q = img + at;
  const int L3 = q[0], L2 = q[st], L1 = q[2 * st], L0 = q[3 * st];
  const int R0 = q[4 * st], R1 = q[5 * st], R2 = q[6 * st], R3 = q[7 * st];
  const int alpha = c_alpha[qp], beta = c_beta[qp];
  if (abs(R0 - L0) >= alpha || abs(R0 - R1) >= beta || abs(L0 - L1) >= beta) return;
  bool strong;
  int c0 = 0;
  if (mode == kIntraSideFilter) {
    const int strength = (int)bs;
    strong = strength == 4;
    if (!strong) c0 = c_tc0[qp][strength];
  } else {
    const double mu = kBsMax * 0.5;
    if (bs > mu) strong = true;
    else if (bs > 0.0) { strong = false; c0 = (int)(bs * (pow2qp - 1)); }
    else return;
  }
  if (strong) {
    const int RL0 = L0 + R0;
    const bool gap = abs(R0 - L0) < ((alpha >> 2) + 2);
    const bool aq = (abs(R0 - R2) < beta) && gap, ap = (abs(L0 - L2) < beta) && gap;
    if (ap) {
      q[3 * st] = (uint8_t)((R1 + ((L1 + RL0) << 1) + L2 + 4) >> 3);
      q[2 * st] = (uint8_t)((L2 + L1 + RL0 + 2) >> 2);
      q[st] = (uint8_t)((((L3 + L2) << 1) + L2 + L1 + RL0 + 4) >> 3);
    } else {
      q[3 * st] = (uint8_t)(((L1 << 1) + L0 + R1 + 2) >> 2);
    }
    if (aq) {
      q[4 * st] = (uint8_t)((L1 + ((R1 + RL0) << 1) + R2 + 4) >> 3);
      q[5 * st] = (uint8_t)((R2 + R0 + L0 + R1 + 2) >> 2);
      q[6 * st] = (uint8_t)((((R3 + R2) << 1) + R2 + R1 + RL0 + 4) >> 3);
    } else {
      q[4 * st] = (uint8_t)(((R1 << 1) + R0 + L1 + 2) >> 2);
    }
  } else {
    const int RL0 = (L0 + R0 + 1) >> 1;
    const bool aq = abs(R0 - R2) < beta, ap = abs(L0 - L2) < beta;
    const int tc = c0 + ap + aq;
    const int dif = clip3(-tc, tc, (((R0 - L0) << 2) + (L1 - R1) + 4) >> 3);
    if (ap) q[2 * st] = (uint8_t)(L1 + clip3(-c0, c0, (L2 + RL0 - (L1 << 1)) >> 1));  // unsigned char +=
    q[3 * st] = (uint8_t)clip3(0, 255, L0 + dif);
    q[4 * st] = (uint8_t)clip3(0, 255, R0 - dif);
    if (aq) q[5 * st] = (uint8_t)(R1 + clip3(-c0, c0, (R2 + RL0 - (R1 << 1)) >> 1));
  }
}

// dir 0: thread = pixel row, its vertical edges left to right.
__global__ void k_deblock_rows(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.y]];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (!f.deblock || r >= f.H) return;
  const int bw = f.W >> 2, i = r >> 2;
  for (int j = 0; j < bw - 1; j++) {
    const int e = i * (bw - 1) + j;
    if (f.dbEdgeMode[e] == kNoFilter) continue;
    deblock_line(f.recons, (size_t)r * f.W + j * 4, 1, f.dbEdgeMode[e], f.dbEdgeBs[e], f.intraQp, f.pow2qp);
  }
}

// dir 1: thread = pixel column, its horizontal edges top to bottom.
__global__ void k_deblock_cols(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.y]];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (!f.deblock || c >= f.W) return;
  const int bw = f.W >> 2, bh = f.H >> 2, nv = (bw - 1) * bh, j = c >> 2;
  for (int i = 0; i < bh - 1; i++) {
    const int e = nv + i * bw + j;
    if (f.dbEdgeMode[e] == kNoFilter) continue;
    deblock_line(f.recons, (size_t)(i * 4) * f.W + c, f.W, f.dbEdgeMode[e], f.dbEdgeBs[e], f.intraQp, f.pow2qp);
  }
}

hipError_t launch_intra(const FrameDev *F, const int *list, int n, int maxBlk88, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_intra_recon, dim3(n), dim3(128), 0, s, F, list);
  hipLaunchKernelGGL(k_intra_si, dim3((maxBlk88 + 255) / 256, n), dim3(256), 0, s, F, list);
  return hipGetLastError();
}

hipError_t launch_deblock(const FrameDev *F, const int *list, int n, int maxW, int maxH, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int maxEdges = 2 * (maxW / 4) * (maxH / 4);
  hipLaunchKernelGGL(k_deblock_decide, dim3((maxEdges + 255) / 256, n), dim3(256), 0, s, F, list);
  hipLaunchKernelGGL(k_deblock_rows, dim3((maxH + 63) / 64, n), dim3(64), 0, s, F, list);
  hipLaunchKernelGGL(k_deblock_cols, dim3((maxW + 63) / 64, n), dim3(64), 0, s, F, list);
  return hipGetLastError();
}

}  // namespace dvc
