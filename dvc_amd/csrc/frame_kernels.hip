// frame_kernels.hip -- side information, transform, correlation-noise-model,
// motion-learning and reconstruction kernels of the WZ frame decode on gfx950.
//
// Each kernel restates one reference function (file:line of DvcDec/src/) with
// the reference's operand order and precision: integer SADs, FP64 motion
// costs and interpolation weights, FP32 noise-model sums in the reference's
// summation order, and the glibc-identical expf of dvc_math.h.  The results
// are bit-identical to the CPU path.  Work is laid out per frame of a batch:
// blockIdx.y (or .z) selects the frame, the other grid dimensions its blocks.
#include <float.h>
#include <algorithm>
#include "dvc_math.h"
#include "frame_kernels.h"
#include "ldpca_kernels.h"
#include "engine_kernels.h"

#pragma clang fp contract(off)

namespace dvc {

// BlockModeFlag value of a WZ block (defines.h:15)
constexpr uint8_t kWzBlock = 0, kSkipBlock = 2;
// Whether 4x4 block b of f is a WZ block (its 8x8 block's mode, as the
// reference's blk88row + (j >> 1) lookup reads it).
__device__ __forceinline__ bool wz_block(const FrameDev &f, int b) {
  if (!f.mode) return true;
  const int bw = f.W >> 2, i = b / bw, j = b - i * bw;
  return f.mode[(i >> 1) * (f.W >> 3) + (j >> 1)] == kWzBlock;
}

#define DVC_K 0.05f  // defines.h:10, promoted to double where it is used

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }

// computeEndPoint (sideInfoCreation.cpp:14-22)
__device__ __forceinline__ int end_point(double st, double vec, double scale, int lower, int upper) {
  int e = (int)((st + vec * scale) + 0.5);
  return e < lower ? lower : (e > upper ? upper : e);
}
// costFunc (sideInfoCreation.cpp:10-12)
__device__ __forceinline__ double cost_func(double mad, double x, double y) {
  return mad * (1.0 + (double)DVC_K * sqrt(x * x + y * y));
}

// Lexicographic (cost, index) minimum across a wave: the first candidate in
// scan order among those of least cost, which is what a strict-< raster scan
// returns.
__device__ __forceinline__ void wave_argmin(double &cost, int &idx) {
  for (int o = 32; o > 0; o >>= 1) {
    double c2 = __shfl_xor(cost, o);
    int i2 = __shfl_xor(idx, o);
    if (c2 < cost || (c2 == cost && i2 < idx)) { cost = c2; idx = i2; }
  }
}

// ------------------------------------------------------------ key frames
// lowPassFilter (filter.cpp:214-233): sum/9 in double equals integer division.
// FIR_filter (filter.cpp:55-117): the 6-tap (1,-5,20,20,-5,1)/32 half-pel
// filter; its dyadic coefficients make (int)(sum + 0.5) then the [0,255]
// clip equal to clip((S + 16) >> 5) (1-D) and clip((S + 512) >> 10) (2-D).
// Thread (i, j) of a grid of rows i < rowsY (rowsY >= H) and columns j:
// low-pass sample (i, j) and the half-pel samples of rows i, i + rowsY, ...
// of the 2H x 2W image; with copyTo set it also copies in[i][j] there.
__device__ __forceinline__ void keyframe_filters_at(const uint8_t *in, uint8_t *lp, uint8_t *half, int W, int H,
                                                    int i, int j, int rowsY, uint8_t *copyTo) {
  const int W2 = 2 * W;
  if (j >= W2) return;
  if (copyTo && i < H && j < W) copyTo[i * W + j] = in[i * W + j];
  if (i < H && j < W) {
    int sum = 0;
    for (int m = i - 1; m <= i + 1; m++) {
      if (m < 0 || m >= H) continue;
      for (int n = j - 1; n <= j + 1; n++) {
        if (n < 0 || n >= W) continue;
        sum += in[m * W + n];
      }
    }
    lp[i * W + j] = (uint8_t)(sum / 9);
  }
  for (int ii = i; ii < 2 * H; ii += rowsY) {
    const int sub = (j & 1) + ((ii & 1) << 1), y = ii >> 1, x = j >> 1;
    const int tap[6] = {1, -5, 20, 20, -5, 1};
    int v;
    if (sub == 0) {
      v = in[imin(H - 1, y) * W + imin(W - 1, x)];
    } else if (sub == 1) {
      int s = 0;
      for (int q = 0; q < 6; q++) s += tap[q] * in[y * W + imax(0, imin(W - 1, x - 2 + q))];
      v = (s + 16) >> 5;
    } else if (sub == 2) {
      int s = 0;
      for (int q = 0; q < 6; q++) s += tap[q] * in[imax(0, imin(H - 1, y - 2 + q)) * W + x];
      v = (s + 16) >> 5;
    } else {
      int s = 0;
      for (int a = 0; a < 6; a++) {
        const int row = imax(0, imin(H - 1, y - 2 + a)) * W;
        int r = 0;
        for (int q = 0; q < 6; q++) r += tap[q] * in[row + imax(0, imin(W - 1, x - 2 + q))];
        s += tap[a] * r;
      }
      v = (s + 512) >> 10;
    }
    half[ii * W2 + j] = (uint8_t)imax(0, imin(255, v));
  }
}

__global__ void k_keyframe_filters(const uint8_t *in, uint8_t *lp, uint8_t *half, int W, int H) {
  keyframe_filters_at(in, lp, half, W, H, blockIdx.y, blockIdx.x * blockDim.x + threadIdx.x, gridDim.y, nullptr);
}

// blockIdx.z = key frame; rows and columns span the largest frame.
__global__ void k_keyframe_batch(const KeyFrameJob *jobs) {
  const KeyFrameJob kj = jobs[blockIdx.z];
  const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * kj.H) return;  // rows in [H, 2H) still own half-pel row i
  keyframe_filters_at(kj.src, kj.lp, kj.half, kj.W, kj.H, i, j, gridDim.y, kj.src != kj.full ? kj.full : nullptr);
}

// ------------------------------------------------------------ forward ME
// forwardME (sideInfoCreation.cpp:61-118): per macroblock of the future
// low-pass frame, exhaustive search of the past low-pass frame over the
// clipped +-32 window; cost SAD*(1 + K*|v|) in double, first minimum in raster
// order.  One workgroup per macroblock.  The window (up to 80 x 80 bytes) is
// staged in LDS as 32-bit words and the macroblock's 64 words sit in scalar
// registers; lane l of every wave takes candidate column l, the four waves
// take every fourth candidate row, and each candidate's SAD is 64 v_sad_u8 on
// window words byte-aligned with v_alignbyte (exact integer SADs, as the
// reference's integer sums are).  The 65th column of a full window is spread
// over the threads afterwards; ties go to the least raster index.
#define FME_LW 21  // staged words per window row: 20 + one zero word
__global__ __launch_bounds__(256) void k_forward_me(const FrameDev *F, const int *list, CostTables ct) {
  const FrameDev &f = F[list[blockIdx.y]];
  const int bw = f.W / 16, nmb = bw * (f.H / 16), mb = blockIdx.x;
  if (mb >= nmb) return;
  const int W = f.W, top = (mb / bw) * 16, left = (mb % bw) * 16;
  const int sT = imax(0, top - 32), sB = imin(f.H - 16, top + 32);
  const int sL = imax(0, left - 32), sR = imin(W - 16, left + 32);  // sL is a multiple of 16
  const int ww = sR - sL + 1, wh = sB - sT + 1, winH = wh + 15, nw = (ww + 15) >> 2;  // nw <= 20
  __shared__ uint32_t win[80 * FME_LW];
  __shared__ uint32_t s_fb[64];
  __shared__ double s_cost[4];
  __shared__ int s_idx[4];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  if (tid < 64) s_fb[tid] = *(const uint32_t *)(f.fut_lp + (size_t)(top + (tid >> 2)) * W + left + 4 * (tid & 3));
  for (int i = tid; i < winH * FME_LW; i += 256) {
    const int r = i / FME_LW, w = i - r * FME_LW;
    win[i] = w < nw ? *(const uint32_t *)(f.past_lp + (size_t)(sT + r) * W + sL + 4 * w) : 0u;
  }
  __syncthreads();
  uint32_t fb[64];
#pragma unroll
  for (int k = 0; k < 64; k++) fb[k] = __builtin_amdgcn_readfirstlane(s_fb[k]);
  auto sad_at = [&](int oy, int ox) -> uint32_t {
    const uint32_t *wr = win + oy * FME_LW + (ox >> 2);
    const int sh = ox & 3;
    uint32_t sad = 0;
#pragma unroll
    for (int m = 0; m < 16; m++)
#pragma unroll
      for (int q = 0; q < 4; q++)
        sad = __builtin_amdgcn_sad_u8(fb[4 * m + q],
                                      __builtin_amdgcn_alignbyte(wr[m * FME_LW + q + 1], wr[m * FME_LW + q], sh), sad);
    return sad;
  };
  double best = DBL_MAX;
  int bidx = 0x7fffffff;
  if (lane < ww) {
    const double *fac = ct.fme + (sL + lane - left + 32);
    for (int oy = wave; oy < wh; oy += 4) {
      const double cost = (double)sad_at(oy, lane) * fac[(sT + oy - top + 32) * 65];
      if (cost < best) { best = cost; bidx = oy * ww + lane; }
    }
  }
  if (ww > 64)
    for (int oy = tid; oy < wh; oy += 256) {
      const double cost = (double)sad_at(oy, 64) * ct.fme[(sT + oy - top + 32) * 65 + sL + 64 - left + 32];
      if (cost < best || (cost == best && oy * ww + 64 < bidx)) { best = cost; bidx = oy * ww + 64; }
    }
  wave_argmin(best, bidx);
  if (lane == 0) { s_cost[wave] = best; s_idx[wave] = bidx; }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < 4; w++)
      if (s_cost[w] < best || (s_cost[w] == best && s_idx[w] < bidx)) { best = s_cost[w]; bidx = s_idx[w]; }
    const int px = sL + bidx % ww, py = sT + bidx / ww;
    MVd r;
    r.px = (double)px; r.py = (double)py; r.fx = 0.0; r.fy = 0.0;
    r.vx = (double)(px - left); r.vy = (double)(py - top);
    f.fwd[mb] = r;
    f.holes[2 * mb] = (double)left * f.iPC + (double)px * f.iFC;
    f.holes[2 * mb + 1] = (double)top * f.iPC + (double)py * f.iFC;
  }
}

// ------------------------------------------------------------ bidirectional ME chain
// One bidirectional search step of bidirectME (sideInfoCreation.cpp:250-306)
// for block blk, run by one wave; the block's range is bounded by its
// neighbours' current MVs (up/left already refined, down/right not yet).
__device__ void bidir_block(const FrameDev &f, MVd *mv, int blk, int bw, int bh, int blockLen, int lane) {
  const int i = blk / bw, j = blk % bw, W2 = 2 * f.W, H2 = 2 * f.H, bl2 = blockLen * 2;
  const int bot = H2 - bl2, right = W2 - bl2, y2 = i * bl2, x2 = j * bl2;
  const int sT = (i - 1 >= 0) ? (int)mv[blk - bw].fy + bl2 : (int)mv[blk].fy;
  const int sB = (i + 1 < bh) ? (int)mv[blk + bw].fy - bl2 : (int)mv[blk].fy;
  const int sL = (j - 1 >= 0) ? (int)mv[blk - 1].fx + bl2 : (int)mv[blk].fx;
  const int sR = (j + 1 < bw) ? (int)mv[blk + 1].fx - bl2 : (int)mv[blk].fx;
  const int nx = sR - sL + 1, ny = sB - sT + 1;
  const double scale = f.iPC / f.iFC;
  double best = DBL_MAX;
  int bidx = 0x7fffffff;
  if (nx > 0 && ny > 0) {
    for (int c = lane; c < nx * ny; c += 64) {
      const int fy = sT + c / nx, fx = sL + c % nx;
      const int py = end_point(y2, (double)(y2 - fy), scale, 0, bot);
      const int px = end_point(x2, (double)(x2 - fx), scale, 0, right);
      int mad = 0;
      for (int m = 0; m < bl2; m += 2) {
        const uint8_t *pr = f.past_half + (py + m) * W2 + px;
        const uint8_t *fr = f.fut_half + (fy + m) * W2 + fx;
        for (int q = 0; q < bl2; q += 2) mad += abs((int)fr[q] - (int)pr[q]);
      }
      const double cost = cost_func((double)mad, (double)(px - fx), (double)(py - fy));
      if (cost < 10000000.0 && (cost < best)) { best = cost; bidx = c; }
    }
  }
  wave_argmin(best, bidx);
  if (lane == 0 && bidx != 0x7fffffff) {
    const int fy = sT + bidx / nx, fx = sL + bidx % nx;
    const int py = end_point(y2, (double)(y2 - fy), scale, 0, bot);
    const int px = end_point(x2, (double)(x2 - fx), scale, 0, right);
    MVd r;
    r.fx = (double)fx; r.fy = (double)fy; r.px = (double)px; r.py = (double)py;
    r.vx = (double)(px - fx); r.vy = (double)(py - fy);
    mv[blk] = r;
  }
}

// Anti-diagonal sweep: blocks on diagonal d = i + j depend only on diagonal
// d-1 (refined) and d+1 (not yet refined), so this order equals the raster
// order of the in-place reference loop.
__device__ void bidir_sweep(const FrameDev &f, MVd *mv, int blockLen) {
  const int bw = f.W / blockLen, bh = f.H / blockLen;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nwaves = blockDim.x >> 6;
  for (int d = 0; d <= bw + bh - 2; d++) {
    const int i0 = imax(0, d - (bw - 1)), i1 = imin(bh - 1, d);
    for (int i = i0 + wave; i <= i1; i += nwaves) bidir_block(f, mv, i * bw + (d - i), bw, bh, blockLen, lane);
    __threadfence_block();
    __syncthreads();
  }
}

// The rest of createSideInfo (sideInfoCreation.cpp:119-143, 510-550) and
// motionFilterAndComp (:312-436) for one frame per workgroup.  With
// fwdOnly set it stops after forwardME's nearest-hole assignment (:119-143)
// and leaves the macroblock MVs unscaled (forwardME's own output).
__global__ __launch_bounds__(1024) void k_si_chain(const FrameDev *F, const int *list, int fwdOnly) {
  const FrameDev &f = F[list[blockIdx.x]];
  const int tid = threadIdx.x, T = blockDim.x, W = f.W, H = f.H;
  const int bw16 = W / 16, nmb = bw16 * (H / 16), bw8 = W / 8, bh8 = H / 8, nb8 = bw8 * bh8;
  const int W2 = 2 * W, H2 = 2 * H;
  // nearest hole per macroblock (:120-142) and the x2 half-pel scaling (:511-521)
  const double maxDst = W * W + H * H;
  for (int i = tid; i < nmb; i += T) {
    const int top = (i / bw16) * 16, left = (i % bw16) * 16;
    double minDst = maxDst;
    int best = 0;
    for (int jj = 0; jj < nmb; jj++) {
      const double hx = f.holes[2 * jj], hy = f.holes[2 * jj + 1];
      const double dst = (left - hx) * (left - hx) + (top - hy) * (top - hy);
      if (dst < minDst) { minDst = dst; best = jj; }
    }
    const MVd b = f.fwd[best];
    MVd o;
    o.fx = (double)end_point(left, -b.vx, f.iFC, 0, W - 16);
    o.fy = (double)end_point(top, -b.vy, f.iFC, 0, H - 16);
    o.px = (double)end_point(left, b.vx, f.iPC, 0, W - 16);
    o.py = (double)end_point(top, b.vy, f.iPC, 0, H - 16);
    if (!fwdOnly) { o.fx *= 2; o.fy *= 2; o.px *= 2; o.py *= 2; }
    o.vx = o.px - o.fx; o.vy = o.py - o.fy;
    f.mbmv[i] = o;
  }
  if (fwdOnly) return;
  __threadfence_block();
  __syncthreads();
  bidir_sweep(f, f.mbmv, 16);
  // 8x8 initial MVs from the 16x16 vectors (:524-545)
  for (int b = tid; b < nb8; b += T) {
    const int i = b / bw8, j = b % bw8, y2 = i << 4, x2 = j << 4;
    const MVd m = f.mbmv[(i >> 1) * bw16 + (j >> 1)];
    MVd r;
    r.fx = (double)end_point(x2, -m.vx, f.iFC, 0, (W - 8) << 1);
    r.fy = (double)end_point(y2, -m.vy, f.iFC, 0, (H - 8) << 1);
    r.px = (double)end_point(x2, m.vx, f.iPC, 0, (W - 8) << 1);
    r.py = (double)end_point(y2, m.vy, f.iPC, 0, (H - 8) << 1);
    r.vx = r.px - r.fx; r.vy = r.py - r.fy;
    f.bmv[b] = r;
  }
  __threadfence_block();
  __syncthreads();
  bidir_sweep(f, f.bmv, 8);
  // MSE per 8x8 block (:324-342); integer-valued sums are exact in double
  for (int b = tid; b < nb8; b += T) {
    const MVd m = f.bmv[b];
    const int fy = (int)m.fy, fx = (int)m.fx, py = (int)m.py, px = (int)m.px;
    double mse = 0.0;
    for (int a = 0; a < 16; a += 2)
      for (int q = 0; q < 16; q += 2) {
        const double d = (double)f.fut_half[(fy + a) * W2 + fx + q] - (double)f.past_half[(py + a) * W2 + px + q];
        mse += d * d;
      }
    f.mse[b] = mse / 64;
  }
  __threadfence_block();
  __syncthreads();
  // weighted vector-median filter over the 3x3 neighbourhood (:345-405)
  for (int b = tid; b < nb8; b += T) {
    const int i = b / bw8, j = b % bw8, y2 = i << 4, x2 = j << 4;
    int nbr[9], cnt = 0;
    for (int m = i - 1; m <= i + 1; m++) {
      if (m < 0 || m >= bh8) continue;
      for (int q = j - 1; q <= j + 1; q++) {
        if (q < 0 || q >= bw8) continue;
        nbr[cnt++] = m * bw8 + q;
      }
    }
    double minW = 0.0;
    int best = nbr[0];
    for (int k = 1; k < cnt; k++) {
      const double dx = f.bmv[best].vx - f.bmv[nbr[k]].vx, dy = f.bmv[best].vy - f.bmv[nbr[k]].vy;
      minW += f.mse[best] / f.mse[nbr[k]] * sqrt(dx * dx + dy * dy);
    }
    for (int c = 1; c < cnt; c++) {
      const int cand = nbr[c];
      double w = 0.0;
      for (int k = 0; k < cnt; k++) {
        if (c == k) continue;
        const double dx = f.bmv[cand].vx - f.bmv[nbr[k]].vx, dy = f.bmv[cand].vy - f.bmv[nbr[k]].vy;
        w += f.mse[cand] / f.mse[nbr[k]] * sqrt(dx * dx + dy * dy);
      }
      if (w < minW) { minW = w; best = cand; }
    }
    const double vx = f.bmv[best].vx, vy = f.bmv[best].vy;
    MVd r;
    r.fx = (double)end_point(x2, -vx, f.iFC, 0, W2 - 16);
    r.fy = (double)end_point(y2, -vy, f.iFC, 0, H2 - 16);
    r.px = (double)end_point(x2, vx, f.iPC, 0, W2 - 16);
    r.py = (double)end_point(y2, vy, f.iPC, 0, H2 - 16);
    r.vx = r.px - r.fx; r.vy = r.py - r.fy;
    f.bmv_tmp[b] = r;
  }
  __threadfence_block();
  __syncthreads();
  // motion compensation (:407-434)
  for (int p = tid; p < W * H; p += T) {
    const int y = p / W, x = p % W, b = (y >> 3) * bw8 + (x >> 3);
    const MVd m = f.bmv_tmp[b];
    const int a = (y & 7) << 1, q = (x & 7) << 1;
    const int fv = f.fut_half[((int)m.fy + a) * W2 + (int)m.fx + q];
    const int pv = f.past_half[((int)m.py + a) * W2 + (int)m.px + q];
    f.si[p] = (uint8_t)(int)((double)fv * f.iFC + (double)pv * f.iPC);
    f.residue[p] = (fv - pv) >> 1;
  }
  __threadfence_block();
  __syncthreads();
  for (int b = tid; b < nb8; b += T) f.bmv[b] = f.bmv_tmp[b];
}

// ------------------------------------------------------------ transform helpers
__device__ __forceinline__ int round_d(double x) { return x >= 0.0 ? (int)(x + 0.5) : -(int)(-x + 0.5); }

// forward44 (transform.cpp:63-104) with the double post-scale of initTransform (:12-33).
__device__ void forward44(int *blk) {
  // b = sqrt(0.4) as the host's correctly rounded double (transform.cpp:14)
  const double a2 = 0.25, ab2 = 0.5 * 0x1.43d136248490fp-1 / 2, b24 = 0.4 / 4;
  int t[16];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int x0 = blk[c], x1 = blk[4 + c], x2 = blk[8 + c], x3 = blk[12 + c];
    t[c] = x0 + x1 + x2 + x3;
    t[4 + c] = (x0 << 1) + x1 - x2 - (x3 << 1);
    t[8 + c] = x0 - x1 - x2 + x3;
    t[12 + c] = x0 - (x1 << 1) + (x2 << 1) - x3;
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int x0 = t[4 * r], x1 = t[4 * r + 1], x2 = t[4 * r + 2], x3 = t[4 * r + 3];
    blk[4 * r] = x0 + x1 + x2 + x3;
    blk[4 * r + 1] = (x0 << 1) + x1 - x2 - (x3 << 1);
    blk[4 * r + 2] = x0 - x1 - x2 + x3;
    blk[4 * r + 3] = x0 - (x1 << 1) + (x2 << 1) - x3;
  }
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = i >> 2, c = i & 3;
    const double ps = ((r & 1) == 0 && (c & 1) == 0) ? a2 : (((r & 1) && (c & 1)) ? b24 : ab2);
    blk[i] = round_d((double)blk[i] * ps);
  }
}

// inverse44 (transform.cpp:106-151) with preScale*64.
__device__ void inverse44(int *blk) {
  const double a2 = 0.25 * 64.0, ab = 0.5 * 0x1.43d136248490fp-1 * 64.0, b2 = 0.4 * 64.0;
  int t[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int r = i >> 2, c = i & 3;
    const double ps = ((r & 1) == 0 && (c & 1) == 0) ? a2 : (((r & 1) && (c & 1)) ? b2 : ab);
    blk[i] = round_d((double)blk[i] * ps);
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int x0 = blk[c], x1 = blk[4 + c], x2 = blk[8 + c], x3 = blk[12 + c];
    t[c] = x0 + x1 + x2 + (x3 >> 1);
    t[4 + c] = x0 + (x1 >> 1) - x2 - x3;
    t[8 + c] = x0 - (x1 >> 1) - x2 + x3;
    t[12 + c] = x0 - x1 + x2 - (x3 >> 1);
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int x0 = t[4 * r], x1 = t[4 * r + 1], x2 = t[4 * r + 2], x3 = t[4 * r + 3];
    blk[4 * r] = x0 + x1 + x2 + (x3 >> 1);
    blk[4 * r + 1] = x0 + (x1 >> 1) - x2 - x3;
    blk[4 * r + 2] = x0 - (x1 >> 1) - x2 + x3;
    blk[4 * r + 3] = x0 - x1 + x2 - (x3 >> 1);
  }
#pragma unroll
  for (int i = 0; i < 16; i++) blk[i] >>= 6;
}

__constant__ int c_zigzag[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
__constant__ int c_qbits[9][16] = {
    {0},
    {4, 3, 0, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {5, 3, 0, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {5, 3, 2, 0, 3, 2, 0, 0, 2, 0, 0, 0, 0, 0, 0, 0},
    {5, 4, 3, 2, 4, 3, 2, 0, 3, 2, 0, 0, 2, 0, 0, 0},
    {5, 4, 3, 2, 4, 3, 2, 2, 3, 2, 2, 0, 2, 2, 0, 0},
    {6, 4, 3, 3, 4, 3, 3, 2, 3, 3, 2, 2, 3, 2, 2, 0},
    {6, 5, 4, 3, 5, 4, 3, 2, 4, 3, 2, 2, 3, 2, 2, 0},
    {7, 6, 5, 4, 6, 5, 4, 3, 5, 4, 3, 2, 4, 3, 2, 0}};

// ------------------------------------------------------------ noise parameter
// forwardTransform of the residue (transform.cpp:153-190) and noiseParameter
// (noiseModel.cpp:9-75) for one frame per workgroup.  The band means and
// second moments are float sums over all blocks in raster order, one band per
// thread, as the reference's loop accumulates them.
__global__ __launch_bounds__(1024) void k_noise_param(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.x]];
  const int tid = threadIdx.x, T = blockDim.x, W = f.W, N = f.N, bw = W / 4;
  __shared__ float s_mean[16], s_var[16];
  __shared__ int s_nbands;
  for (int b = tid; b < N; b += T) {
    const int top = (b / bw) * 4, left = (b % bw) * 4;
    int blk[16];
    for (int m = 0; m < 4; m++)
      for (int n = 0; n < 4; n++) blk[4 * m + n] = f.residue[(top + m) * W + left + n];
    forward44(blk);
    for (int m = 0; m < 4; m++)
      for (int n = 0; n < 4; n++) f.sideTrans[(top + m) * W + left + n] = blk[4 * m + n];
  }
  if (tid == 0) {
    int nb = 0;
    while (nb < 16 && c_qbits[f.qindex][c_zigzag[nb]] > 0) nb++;
    s_nbands = nb;
  }
  __threadfence_block();
  __syncthreads();
  const int nb = s_nbands;
  if (tid < nb) {
    const int index = c_zigzag[tid];
    float mean = 0.0f, var = 0.0f;
    for (int b = 0; b < N; b++) {
      const int top = (b / bw) * 4, left = (b % bw) * 4;
      const float c = fabsf((float)f.sideTrans[(top + (index >> 2)) * W + left + (index & 3)]);
      mean += c;
      var += c * c;
    }
    mean /= (float)N;
    var = var / (float)N - mean * mean;
    s_mean[tid] = mean;
    s_var[tid] = var;
  }
  __syncthreads();
  for (int e = tid; e < N * nb; e += T) {
    const int b = e / nb, band = e % nb, index = c_zigzag[band];
    const int top = (b / bw) * 4, left = (b % bw) * 4;
    const int ci = (top + (index >> 2)) * W + left + (index & 3);
    const float c = fabsf((float)f.sideTrans[ci]);
    const float dst = (float)(int)c - s_mean[band];
    const float dst2 = dst * dst;
    f.alpha[ci] = dst2 > s_var[band] ? sqrtf(2 / dst2) : sqrtf(2 / s_var[band]);
  }
}

// ------------------------------------------------------------ over-complete transform
// overCompleteTransform (transform.cpp:275-298); with initSideTrans, also
// init_si_ML (motionLearning.cpp:49-68) from the block-aligned positions.
// Each position's coefficient of the band whose noise distribution comes next
// (band 0 at the frame's start, else band + 1) also goes to the octBand plane,
// so that kernel reads 4 contiguous bytes per window position instead of one
// int out of each 64-byte record.
// A workgroup's 256 consecutive positions span at most 3 position rows of a
// frame at least 176 wide, so the side-information rows under them (at most
// 6) are staged in LDS as words with aligned loads, and each position's 4x4
// pixels are taken from there (4 byte alignments) instead of 16 byte loads.
// The 64-byte records of a wave's 64 positions go out through LDS so that each
// store instruction writes 1 KB of consecutive bytes.
#define OCT_ROWS 6
#define OCT_LW 89  // words per staged row: 352 / 4 + one for the byte alignment
__global__ __launch_bounds__(256) void k_oct(const FrameDev *F, const int *list, int initSideTrans) {
  const FrameDev &f = F[list[blockIdx.y]];
  const int ow = f.W - 3, npos = ow * (f.H - 3), W = f.W, nw = W >> 2;
  const int p0 = blockIdx.x * 256;
  if (p0 >= npos) return;  // uniform over the workgroup
  __shared__ uint32_t s_rows[OCT_ROWS * OCT_LW];
  __shared__ int4 s_rec[256 * 4];  // the workgroup's records, position-major
  const int top0 = p0 / ow, rows = imin(f.H, (imin(p0 + 255, npos - 1) / ow) + 4) - top0;  // <= OCT_ROWS
  for (int i = threadIdx.x; i < rows * OCT_LW; i += 256) {
    const int r = i / OCT_LW, w = i - r * OCT_LW;
    s_rows[i] = w < nw ? *(const uint32_t *)(f.si + (size_t)(top0 + r) * W + 4 * w) : 0u;
  }
  __syncthreads();
  const int p = p0 + threadIdx.x;
  const bool valid = p < npos;
  const int top = valid ? p / ow : top0, left = valid ? p % ow : 0, j = left >> 2, sh = left & 3;
  const uint32_t *R = s_rows + (top - top0) * OCT_LW + j;
  int blk[16];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const uint32_t v = __builtin_amdgcn_alignbyte(R[m * OCT_LW + 1], R[m * OCT_LW], sh);
#pragma unroll
    for (int n = 0; n < 4; n++) blk[4 * m + n] = (v >> (8 * n)) & 0xff;
  }
  forward44(blk);
  int4 *rec = s_rec + 4 * threadIdx.x;
  rec[0] = make_int4(blk[0], blk[1], blk[2], blk[3]);
  rec[1] = make_int4(blk[4], blk[5], blk[6], blk[7]);
  rec[2] = make_int4(blk[8], blk[9], blk[10], blk[11]);
  rec[3] = make_int4(blk[12], blk[13], blk[14], blk[15]);
  if (valid) {
    const int nb = initSideTrans ? 0 : f.band + 1;
    if (nb < 16) f.octBand[p] = blk[c_zigzag[nb]];
    if (initSideTrans && (top & 3) == 0 && (left & 3) == 0 && top <= f.H - 4 && left <= f.W - 4)
      for (int m = 0; m < 4; m++)
        for (int n = 0; n < 4; n++) f.sideTrans[(top + m) * W + left + n] = blk[4 * m + n];
  }
  __syncthreads();
  // store i writes the workgroup's 16-byte pieces 256 i + t: position p0 + 64 i + t / 4
  int4 *dst = (int4 *)(f.oct + (size_t)16 * p0);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int piece = 256 * i + threadIdx.x;
    if (p0 + (piece >> 2) < npos) dst[piece] = s_rec[piece];
  }
}

// octBand from a complete over-complete transform (coefficient `index`).
__global__ __launch_bounds__(256) void k_oct_plane(const FrameDev *F, const int *list, int index) {
  const FrameDev &f = F[list[blockIdx.y]];
  const int npos = (f.W - 3) * (f.H - 3);
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npos) f.octBand[p] = f.oct[(size_t)16 * p + index];
}

// SMF window of WZ block b (initSMF, motionLearning.cpp:6-47).
struct SmfWin { int top, left, sT, sB, sL, sR, w, len, self; };
__device__ __forceinline__ SmfWin smf_window(const FrameDev &f, int b) {
  SmfWin s;
  const int bw = f.W / 4;
  s.top = (b / bw) * 4; s.left = (b % bw) * 4;
  s.sT = imax(0, s.top - 4); s.sB = imin(f.H - 4, s.top + 4);
  s.sL = imax(0, s.left - 4); s.sR = imin(f.W - 4, s.left + 4);
  s.w = s.sR - s.sL + 1;
  s.len = s.w * (s.sB - s.sT + 1);
  s.self = (s.top - s.sT) * s.w + (s.left - s.sL);
  return s;
}

__global__ void k_smf_init(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.y]];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= f.N) return;
  const SmfWin s = smf_window(f, b);
  for (int p = 0; p < 81; p++) f.smf[(size_t)81 * b + p] = (p == s.self) ? 1.0f : 0.0f;
}

// ------------------------------------------------------------ noise distribution
// noiseDistribution (noiseModel.cpp:77-132): DC prob[k] = expf(-a*|k - SI_DC|);
// AC prob[k] = sum_m p_m * expf(-a*|k - range - SI_m|), m in window raster
// order.  A window term whose probability is exactly +0 adds +0 to a
// non-negative sum and is skipped when a is finite.  The exponent depends on
// the integer distance d = |k - range - SI_m| only, so each workgroup first
// tabulates expf(-a*(float)d) for the distances its block can reach (the same
// values, bit for bit) and the mixture sums then read the table.  One wave
// per (4x4 block, frame): a block has 2*range+1 <= ~520 values at most
// Qindex 8, and the small workgroup with a 4 KB table lets a CU hold many
// blocks at once to hide the LDS and global latencies of each one's phases.
#define ND_TAB 1024  // table entries (4 KB of LDS); larger distances call expf directly
#define ND_T 64      // threads per block distribution (one wave)
// V values k + i*ND_T (i < V) of one block's distribution: sum over the window
// terms m in order of p_m * tab[value + off_m].  Values at or beyond rmax are
// computed from the table's padding (never from outside it) and not stored.
template <int V>
__device__ __forceinline__ void nd_values(const float *tab, const int2 *ps, int n, int k, int rmax, float *out) {
  float acc[V];
#pragma unroll
  for (int i = 0; i < V; i++) acc[i] = 0.0f;
  for (int m = 0; m < n; m++) {
    const int2 t = ps[m];
    const float pr = __int_as_float(t.x);
    const float *tk = tab + k + t.y;
#pragma unroll
    for (int i = 0; i < V; i++) acc[i] += pr * tk[i * ND_T];
  }
#pragma unroll
  for (int i = 0; i < V; i++)
    if (k + i * ND_T < rmax) out[k + i * ND_T] = acc[i];
}
__global__ __launch_bounds__(64) void k_noise_dist(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.y]];
  const int b = blockIdx.x;
  if (b >= f.N || !wz_block(f, b)) return;  // only SIR candidates have a distribution
  const int tid = threadIdx.x, T = ND_T;
  const int ow = f.W - 3, index = f.index, range = f.range;
  const SmfWin s = smf_window(f, b);
  const float a = f.alpha[(s.top + (index >> 2)) * f.W + s.left + (index & 3)];
  float *out = f.dist + (size_t)b * f.distStride;
  if (index == 0) {
    const int side = f.octBand[s.top * ow + s.left];
    for (int k = tid; k < range; k += T) out[k] = dvc_expf(-a * (float)abs(k - side));
    return;
  }
  __shared__ int2 s_ps[81];  // (probability bits, SI coefficient) of the kept window terms
  __shared__ int s_n, s_maxabs;
  // table entries [0, 2D] with 2D + 1 <= ND_TAB, and 3 * ND_T of padding that a
  // lane's values beyond rmax (at most 3 * ND_T further on) may read
  __shared__ float s_tab[ND_TAB + 3 * ND_T];
  if (tid < 64) {
    // wave 0 compacts the window's non-zero terms in raster order
    const bool fin = isfinite(a);
    int n = 0, mx = 0;
    for (int base = 0; base < s.len; base += 64) {
      const int p = base + tid;
      float pr = 0.0f;
      bool keep = false;
      if (p < s.len) {
        pr = f.smf[(size_t)81 * b + p];
        keep = !(fin && pr == 0.0f);
      }
      const unsigned long long mask = __ballot(keep);
      if (keep) {
        const int pos = n + __popcll(mask & ((1ull << tid) - 1ull));
        const int m = s.sT + p / s.w, q = s.sL + p % s.w;
        const int sd = f.octBand[m * ow + q];
        s_ps[pos] = make_int2(__float_as_int(pr), sd);
        mx = imax(mx, abs(sd));
      }
      n += __popcll(mask);
    }
    for (int o = 32; o > 0; o >>= 1) mx = imax(mx, __shfl_xor(mx, o));
    if (tid == 0) { s_n = n; s_maxabs = mx; }
  }
  __syncthreads();
  const int n = s_n;
  // |k - range - SI_m| <= range + max|SI_m| = D for k in [0, 2*range]
  const int D = range + s_maxabs;
  const int rmax = 2 * range + 1;
  if (2 * D + 1 <= ND_TAB) {
    // two-sided table s_tab[D + t] = expf(-a*|t|), t in [-D, D]: term m of
    // value k reads s_tab[k + off_m] with off_m = D - range - SI_m
    for (int d = tid; d <= D; d += T) {
      const float v = dvc_expf(-a * (float)d);
      s_tab[D + d] = v;
      s_tab[D - d] = v;
    }
    for (int m = tid; m < n; m += T) s_ps[m].y = D - range - s_ps[m].y;
    __syncthreads();
    // a lane's values k, k + T, ... share each term's offset and carry
    // independent sums; each value still adds its terms in window order.  The
    // number of values per lane is the same on every lane (groups of up to 4).
    int k = tid, left = (rmax + T - 1) / T;
    for (; left >= 4; left -= 4, k += 4 * T) nd_values<4>(s_tab, s_ps, n, k, rmax, out);
    if (left == 3) nd_values<3>(s_tab, s_ps, n, k, rmax, out);
    else if (left == 2) nd_values<2>(s_tab, s_ps, n, k, rmax, out);
    else if (left == 1) nd_values<1>(s_tab, s_ps, n, k, rmax, out);
  } else {
    const int need = D + 1;
    const int tl = imin(need, ND_TAB);
    for (int d = tid; d < tl; d += T) s_tab[d] = dvc_expf(-a * (float)d);
    __syncthreads();
    for (int k = tid; k < rmax; k += T) {
      const int pc = k - range;
      float sum = 0.0f;
      for (int m = 0; m < n; m++) {
        const int2 ps = s_ps[m];
        const int d = abs(pc - ps.y);
        const float e = d < tl ? s_tab[d] : dvc_expf(-a * (float)d);
        sum += __int_as_float(ps.x) * e;
      }
      out[k] = sum;
    }
  }
}

// ------------------------------------------------------------ conditional bit probability
// condBitProb (noiseModel.cpp:191-310): interval sums in ascending k, p0/(p0+p1),
// clipped likelihood ratio.  One wave per 64 blocks: the wave stages the span
// of each block's two intervals through LDS in chunks of 32 values (each load
// a coalesced 128-byte piece of one block's distribution row), and each lane
// then adds its own block's values in ascending k, as the reference's loops do.
__global__ __launch_bounds__(64) void k_condbit(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.y]];
  const int lane = threadIdx.x;
  const int b0 = blockIdx.x * 64;
  if (b0 >= f.N) return;  // uniform over the wave
  __shared__ float s_v[64 * 33];
  __shared__ int s_lo[64], s_hi[64];
  const int b = b0 + lane;
  const bool valid = b < f.N;
  const bool wz = valid && wz_block(f, b);  // condBitProb walks the SIR's WZ blocks
  int v0l = 0, v0u = 0, v1l = 0, v1u = 0;
  if (wz) {
    const int bw = f.W / 4, top = (b / bw) * 4, left = (b % bw) * 4;
    const int index = f.index, bit = f.bit, step = f.step, range = f.range;
    const int iq = (1 << (bit - 1)) - 1;
    const int rc = f.reconsTrans[(top + (index >> 2)) * f.W + left + (index & 3)];
    int lim;
    if (index == 0) {
      const int q0l = rc << bit, q1l = ((rc << 1) + 1) << (bit - 1);
      const int q0u = q0l + iq, q1u = q1l + iq;
      v0l = q0l * step; v1l = q1l * step; v0u = (q0u + 1) * step; v1u = (q1u + 1) * step;
      lim = range;
    } else {
      const int ls = f.levelShift;
      const int q0l = (rc << bit) - ls, q1l = (((rc << 1) + 1) << (bit - 1)) - ls;
      const int q0u = q0l + iq;
      int q1u = q1l + iq;
      if (q1u > ls) q1u = ls;
      v0l = (q0l <= 0) ? (q0l - 1) * step + 1 : q0l * step;
      v1l = (q1l <= 0) ? (q1l - 1) * step + 1 : q1l * step;
      v0u = (q0u < 0) ? q0u * step + 1 : (q0u + 1) * step;
      v1u = (q1u < 0) ? q1u * step + 1 : (q1u + 1) * step;
      v0l += range; v0u += range; v1l += range; v1u += range;
      lim = 2 * range + 1;
    }
    v0l = imax(v0l, 0); v1l = imax(v1l, 0); v0u = imin(v0u, lim); v1u = imin(v1u, lim);
  }
  // the span read for this block: [lo, hi) covers both (non-empty) intervals
  int lo = 0, hi = 0;
  if (v0u > v0l) { lo = v0l; hi = v0u; }
  if (v1u > v1l) {
    lo = hi > lo ? imin(lo, v1l) : v1l;
    hi = imax(hi, v1u);
  }
  s_lo[lane] = lo;
  s_hi[lane] = hi;
  int span = hi - lo;
  for (int o = 32; o > 0; o >>= 1) span = imax(span, __shfl_xor(span, o));
  __syncthreads();
  const size_t stride = f.distStride;
  float p0 = 0.0f, p1 = 0.0f;
  // chunk c0 of every block's span, 32 values per block: lane (j parity,
  // i) fetches value i of blocks jj + (lane >> 5); the next chunk's loads are
  // issued before this chunk is summed
  float nxt[32];
  auto fetch = [&](int c0) {
#pragma unroll
    for (int jj = 0; jj < 64; jj += 2) {
      const int j = jj + (lane >> 5), i = lane & 31;
      const int k = s_lo[j] + c0 + i;
      nxt[jj >> 1] = (b0 + j < f.N && k < s_hi[j]) ? f.dist[(size_t)(b0 + j) * stride + k] : 0.0f;
    }
  };
  if (span > 0) fetch(0);
  for (int c0 = 0; c0 < span; c0 += 32) {
#pragma unroll
    for (int jj = 0; jj < 64; jj += 2) s_v[(jj + (lane >> 5)) * 33 + (lane & 31)] = nxt[jj >> 1];
    __syncthreads();
    if (c0 + 32 < span) fetch(c0 + 32);
    for (int i = 0; i < 32; i++) {
      const int k = lo + c0 + i;
      const float v = s_v[lane * 33 + i];
      if (k >= v0l && k < v0u) p0 += v;
      if (k >= v1l && k < v1u) p1 += v;
    }
    __syncthreads();
  }
  if (!valid) return;
  if (!wz) {
    // the frame's LR are first set to 0.9999f / 0.0001f when it has skip or
    // intra blocks (DVC_Decoder.cpp:203-207); WZ blocks overwrite theirs
    f.LR[b] = 0.9999f / 0.0001f;
    return;
  }
  p0 = p0 / (p0 + p1);
  p1 = 1 - p0;
  float lr;
  if ((double)p0 <= 0.0001) lr = (float)(0.0001 / 0.9999);
  else if ((double)p0 >= 0.9999) lr = (float)(0.9999 / 0.0001);
  else lr = p0 / p1;
  f.LR[b] = lr;
}

// ------------------------------------------------------------ bitplane -> coefficients
// putBitplaneToCoeff (bitplane.cpp:4-22) with the bits of the code the rate
// search accepted (its record names the decision slot).
__global__ void k_put_bitplane(const FrameDev *F, const int *list, const BpFrameLink *links) {
  const int fi = list[blockIdx.y];
  const FrameDev &f = F[fi];
  const BpFrameLink &lk = links[blockIdx.y];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= f.N) return;
  const int best = lk.state->best;
  const int slot = (best >= 0 && best < kMaxCodes) ? lk.records[best].z : -1;
  if (slot < 0 || slot >= lk.slot_capacity) return;  // the host reports the missing record
  const uint8_t bitv = lk.slots[(size_t)slot * f.N + b];
  f.decoded[b] = bitv;
  if (lk.carry) lk.carry[b] = bitv;  // ParallelMode 1: the next bitplane starts from these decisions
  if (!wz_block(f, b)) return;      // putBitplaneToCoeff (bitplane.cpp:4-22) skips other blocks
  const int bw = f.W / 4, top = (b / bw) * 4, left = (b % bw) * 4, index = f.index;
  int *p = &f.reconsTrans[(top + (index >> 2)) * f.W + left + (index & 3)];
  *p = (*p << 1) + (int)bitv;
}

// ------------------------------------------------------------ reconstruction
// reconstruction (reconstruction.cpp:184-289).  exp is dvc_exp, the host C
// library's double exp restated bit for bit, and the remaining operations are
// the same IEEE double operations in the same order without contraction, so
// every reconstructed coefficient equals the CPU path's.
__global__ void k_reconstruct(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.y]];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= f.N || !wz_block(f, b)) return;
  const int bw = f.W / 4, top = (b / bw) * 4, left = (b % bw) * 4, index = f.index, step = f.step;
  const int ci = (top + (index >> 2)) * f.W + left + (index & 3);
  const int side = f.sideTrans[ci];
  const double a = f.alpha[ci];
  int lo, hi;
  if (index == 0) {
    const int q = f.reconsTrans[ci];
    lo = q * step; hi = (q + 1) * step;
  } else {
    const int q = f.reconsTrans[ci] - f.levelShift;
    if (q == 0) { lo = -step + 1; hi = step; }
    else if (q > 0) { lo = q * step; hi = (q + 1) * step; }
    else { lo = (q - 1) * step + 1; hi = q * step + 1; }
  }
  int v;
  if (a <= 0.0001) {
    v = (lo + hi) / 2;
  } else if (side < lo) {
    v = lo + round_d(1.0 / a + (double)step / (1.0 - dvc_exp(a * (double)step)));
  } else if (side >= hi) {
    v = hi - round_d(1.0 / a + (double)step / (1.0 - dvc_exp(a * (double)step)));
  } else {
    const double r = side - lo, d = hi - side;
    const double er = dvc_exp(-a * r), ed = dvc_exp(-a * d);
    v = side + round_d(((r + 1.0 / a) * er - (d + 1.0 / a) * ed) / (2.0 - (er + ed)));
  }
  f.sideTrans[ci] = v;
  if (f.dbFlag) f.dbFlag[ci] = v != side;  // deblockFlag: reconstruction moved the coefficient
}

// inverseTransform (transform.cpp:234-273).
__global__ void k_inverse_transform(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.y]];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= f.N || !wz_block(f, b)) return;  // the SIR lists WZ blocks only
  const int bw = f.W / 4, top = (b / bw) * 4, left = (b % bw) * 4, W = f.W;
  int blk[16];
  for (int m = 0; m < 4; m++)
    for (int n = 0; n < 4; n++) blk[4 * m + n] = f.sideTrans[(top + m) * W + left + n];
  inverse44(blk);
  for (int m = 0; m < 4; m++)
    for (int n = 0; n < 4; n++) {
      const int c = blk[4 * m + n];
      f.recons[(top + m) * W + left + n] = (uint8_t)(c < 0 ? 0 : (c > 255 ? 255 : c));
    }
}

// ------------------------------------------------------------ SI refresh
// updateSI + subPelRefine (sideInfoCreation.cpp:598-736): per WZ 4x4 block of
// the partially reconstructed frame, separate full-pel argmins on the past
// and future key frames (cost SAD*(1+K*|2v|) in double), 9-position half-pel
// refinement on each interpolated frame, then SI = (int)(P*iPC + F*iFC).
// One wave per block, 4 horizontally adjacent blocks per workgroup (the frame
// width is a multiple of 16, so the 4 share a block row).  The union of their
// search windows (up to 68 rows x 80 bytes of each key frame) is staged in LDS
// as 32-bit words, each wave staging every fourth row of both frames, the two
// frames' rows interleaved so that one address reaches both; lane l
// sweeps candidate column sL+l down the window, forming each candidate's
// 4-pixel rows with v_alignbyte and its SAD with v_sad_u8 (exact integer SADs,
// as the reference's double sums of integer differences are).  A lane keeps
// its best cost and the row it was found on (the row is wave-uniform, so the
// select takes it from a scalar register); the rows above the block and the
// rows from the block down run as two loops, so the cost factor's LDS address
// steps by a constant.  The 65th column of a full window is spread over the
// lanes afterwards.  Ties resolve to the first candidate in raster order: the
// wave's minimum cost first, then the least (row, column) key among the lanes
// that hold it.
// A cost (double)sad * fac is formed as fma(2^52 + sad, fac, -2^52 * fac): the
// double with high word 0x43300000 and low word sad is 2^52 + sad exactly, the
// fused product-sum is sad * fac before its one rounding, and -2^52 * fac is
// exact, so the result has the bits of the rounded product (sad = 0 gives +0
// both ways) without a conversion.
#define USI_LW 21   // staged words per window row: 20 + one zero word
#define USI_ROWS 68
__device__ __forceinline__ uint32_t usi_word(const uint32_t *R, int r, int j, int sh) {
  return __builtin_amdgcn_alignbyte(R[r * 2 * USI_LW + j + 1], R[r * 2 * USI_LW + j], sh);
}
__device__ __forceinline__ uint32_t sad4(const uint32_t *bk, uint32_t r0, uint32_t r1, uint32_t r2, uint32_t r3) {
  uint32_t s = __builtin_amdgcn_sad_u8(bk[0], r0, 0u);
  s = __builtin_amdgcn_sad_u8(bk[1], r1, s);
  s = __builtin_amdgcn_sad_u8(bk[2], r2, s);
  return __builtin_amdgcn_sad_u8(bk[3], r3, s);
}
// v_min_f64 of two values known to be numbers (the compiler would otherwise
// canonicalise a loop-carried operand first)
__device__ __forceinline__ double dmin_num(double a, double b) {
  double r;
  __asm__("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ void argmin_take(double c, int i, double &best, int &bi) {
  if (c < best || (c == best && i < bi)) { best = c; bi = i; }
}
// (cost, key) -> the wave's least cost and, among the lanes holding it, the
// least key; every lane gets both.  Costs are finite or DBL_MAX, never NaN.
__device__ __forceinline__ void wave_argmin_key(double &cost, int &key) {
  double m = cost;
  for (int o = 32; o > 0; o >>= 1) m = dmin_num(m, __shfl_xor(m, o));
  int k = cost == m ? key : 0x7fffffff;
  for (int o = 32; o > 0; o >>= 1) k = imin(k, __shfl_xor(k, o));
  cost = m;
  key = k;
}
__global__ __launch_bounds__(256) void k_update_si(const FrameDev *F, const int *list, CostTables ct) {
  const FrameDev &f = F[list[blockIdx.y]];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int b0 = blockIdx.x * 4;
  if (b0 >= f.N) return;  // uniform over the workgroup
  const int W = f.W, H = f.H, W2 = 2 * W, H2 = 2 * H, bw = W / 4;
  __shared__ uint32_t s_win[USI_ROWS * 2 * USI_LW];  // [row][past, future][word]
  __shared__ double s_fac[33 * 33];  // by (|dy|, |dx|)
  // union window of the 4 blocks: rows [wT, wB], byte columns [cL, cR]
  const int row0 = (b0 / bw) * 4, left0 = (b0 % bw) * 4;
  const int wT = imax(0, row0 - 32), wB = imin(H - 4, row0 + 32) + 3;
  const int cL = imax(0, left0 - 32) & ~3, cR = imin(W - 4, left0 + 12 + 32) + 3;
  const int nrows = wB - wT + 1, nw = (cR + 1 - cL) >> 2;  // nrows <= 68, nw <= 20
  {
    // lanes 0..20 stage a past-frame row's words, lanes 32..52 the future's
    const int half = lane >> 5, w = lane & 31;
    if (w < USI_LW) {
      const uint8_t *src = half ? f.fut : f.past;
      uint32_t *dst = s_win + half * USI_LW;
      for (int r = wave; r < nrows; r += 4)
        dst[r * 2 * USI_LW + w] = w < nw ? *(const uint32_t *)(src + (size_t)(wT + r) * W + cL + 4 * w) : 0u;
    }
    for (int dy = wave; dy < 33; dy += 4)
      if (lane < 33) s_fac[dy * 33 + lane] = ct.usi[(dy + 32) * 65 + lane + 32];
  }
  __syncthreads();
  const int b = b0 + wave;
  if (b >= f.N || !wz_block(f, b)) return;  // updateSI walks the SIR's WZ blocks
  const int top = row0, left = (b % bw) * 4;
  const int sT = imax(0, top - 32), sB = imin(H - 4, top + 32), sL = imax(0, left - 32), sR = imin(W - 4, left + 32);
  const int ww = sR - sL + 1, nr = sB - sT + 1;
  uint32_t bk[4];
#pragma unroll
  for (int m = 0; m < 4; m++) bk[m] = *(const uint32_t *)(f.recons + (size_t)(top + m) * W + left);
  const uint32_t *P = s_win, *Fu = s_win + USI_LW;
  double pBest = DBL_MAX, fBest = DBL_MAX;
  int pKey = 0x7fffffff, fKey = 0x7fffffff;  // (row - sT) * 128 + column - sL
  if (lane < ww) {
    const int x = sL + lane, xo = x - cL, j = xo >> 2, sh = xo & 3;
    const double *facCol = s_fac + abs(x - left);
    int r = sT - wT, pRow = 0, fRow = 0;
    uint32_t p0 = usi_word(P, r, j, sh), p1 = usi_word(P, r + 1, j, sh), p2 = usi_word(P, r + 2, j, sh);
    uint32_t f0 = usi_word(Fu, r, j, sh), f1 = usi_word(Fu, r + 1, j, sh), f2 = usi_word(Fu, r + 2, j, sh);
    // the high words of 2^52 + sad, one register per frame so that each SAD
    // can land in the low half of its own pair
    int hiP = 0x43300000, hiF = 0x43300000;
    __asm__ volatile("" : "+v"(hiP));
    __asm__ volatile("" : "+v"(hiF));
    auto step = [&](int yr, double fac) {
      const uint32_t p3 = usi_word(P, r + 3, j, sh), f3 = usi_word(Fu, r + 3, j, sh);
      const uint32_t ps = sad4(bk, p0, p1, p2, p3), fs = sad4(bk, f0, f1, f2, f3);
      const double facn = fac * -4503599627370496.0;  // -2^52 fac
      const double pc = fma(__hiloint2double(hiP, (int)ps), fac, facn);
      const double fc = fma(__hiloint2double(hiF, (int)fs), fac, facn);
      // costs are positive and never NaN: the running minimum is a plain
      // v_min_f64, and only the row needs the strict comparison
      pRow = pc < pBest ? yr : pRow;
      fRow = fc < fBest ? yr : fRow;
      pBest = dmin_num(pBest, pc);
      fBest = dmin_num(fBest, fc);
      p0 = p1; p1 = p2; p2 = p3;
      f0 = f1; f1 = f2; f2 = f3;
      r++;
    };
    int y = sT;
    // rows above the block: |y - top| falls, so four rows at a time read the
    // table upwards from the lowest of their entries (LDS offsets are positive)
    const double *lo = facCol + 33 * (top - sT - 3);
    for (; y + 4 <= top; y += 4, lo -= 4 * 33) {
      step(y - sT, lo[99]);
      step(y + 1 - sT, lo[66]);
      step(y + 2 - sT, lo[33]);
      step(y + 3 - sT, lo[0]);
    }
    for (; y < top; y++) step(y - sT, facCol[33 * (top - y)]);
    const double *fp = facCol;  // rows from the block down: |y - top| rises
    for (; y + 4 <= sB + 1; y += 4, fp += 4 * 33) {
      step(y - sT, fp[0]);
      step(y + 1 - sT, fp[33]);
      step(y + 2 - sT, fp[66]);
      step(y + 3 - sT, fp[99]);
    }
    for (; y <= sB; y++, fp += 33) step(y - sT, *fp);
    pKey = pRow * 128 + lane;
    fKey = fRow * 128 + lane;
  }
  if (ww > 64) {
    // column sL+64: one candidate row per lane
    const int x = sL + 64, xo = x - cL, j = xo >> 2, sh = xo & 3;
    for (int i = lane; i < nr; i += 64) {
      const int y = sT + i, r = y - wT;
      const uint32_t ps = sad4(bk, usi_word(P, r, j, sh), usi_word(P, r + 1, j, sh), usi_word(P, r + 2, j, sh),
                               usi_word(P, r + 3, j, sh));
      const uint32_t fs = sad4(bk, usi_word(Fu, r, j, sh), usi_word(Fu, r + 1, j, sh), usi_word(Fu, r + 2, j, sh),
                               usi_word(Fu, r + 3, j, sh));
      const double fac = s_fac[33 * abs(y - top) + abs(x - left)];
      argmin_take((double)ps * fac, i * 128 + 64, pBest, pKey);
      argmin_take((double)fs * fac, i * 128 + 64, fBest, fKey);
    }
  }
  wave_argmin_key(pBest, pKey);
  wave_argmin_key(fBest, fKey);
  const int pbx = sL + (pKey & 127), pby = sT + (pKey >> 7), fbx = sL + (fKey & 127), fby = sT + (fKey >> 7);
  // subPelRefine: lanes 0-8 evaluate the 9 positions of the past frame, 16-24 the future
  const int which = lane >> 4, pos = lane & 15;
  double cost = DBL_MAX;
  int cidx = 0x7fffffff;
  if (which < 2 && pos < 9) {
    const int bx = which ? fbx : pbx, by = which ? fby : pby;
    const uint8_t *ref = which ? f.fut_half : f.past_half;
    const int t = 2 * by + pos / 3 - 1, l = 2 * bx + pos % 3 - 1;
    if (t >= 0 && t <= H2 - 8 && l >= 0 && l <= W2 - 8) {
      // the 4 samples of block row m sit 2 bytes apart in the half-pel frame
      uint32_t s = 0;
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const uint8_t *rp = ref + (size_t)(t + 2 * m) * W2 + l;
        const uint32_t row = (uint32_t)rp[0] | (uint32_t)rp[2] << 8 | (uint32_t)rp[4] << 16 | (uint32_t)rp[6] << 24;
        s = __builtin_amdgcn_sad_u8(bk[m], row, s);
      }
      cost = (double)s;
      cidx = pos;
    }
  }
  // argmin within each 16-lane group (first minimum in (i, j) order)
  for (int o = 8; o > 0; o >>= 1) {
    const double c2 = __shfl_xor(cost, o);
    const int i2 = __shfl_xor(cidx, o);
    if (c2 < cost || (c2 == cost && i2 < cidx)) { cost = c2; cidx = i2; }
  }
  const int pSub = __shfl(cidx, 0), fSub = __shfl(cidx, 16);
  const int phx = 2 * pbx + pSub % 3 - 1, phy = 2 * pby + pSub / 3 - 1;
  const int fhx = 2 * fbx + fSub % 3 - 1, fhy = 2 * fby + fSub / 3 - 1;
  if (lane < 16) {
    const int m = lane >> 2, n = lane & 3;
    const int pv = f.past_half[(phy + 2 * m) * W2 + phx + 2 * n];
    const int fv = f.fut_half[(fhy + 2 * m) * W2 + fhx + 2 * n];
    f.si[(top + m) * W + left + n] = (uint8_t)(int)((double)pv * f.iPC + (double)fv * f.iFC);
  }
}

// ------------------------------------------------------------ motion learning
// updateSMF (motionLearning.cpp:70-127): p = expf(-MU*SSE) over the +-4
// window, normalised by the float sum in raster order.  One wave per 4x4
// block.  Every pixel difference is an integer of magnitude at most 255 and
// every partial SSE stays below 2^24, so the reference's float SSE is exact in
// any order and equals the integer sum of squares; it is formed here as
// sum(a^2) + sum(b^2) - 2 sum(a*b) with v_dot4_u32_u8 on 4-pixel rows (a the
// reconstructed block, b the side information under the window position),
// reading b from the block's side-information window (at most 12 rows x 12
// bytes) staged in LDS as words.  Lane l evaluates window positions l and
// l+64; the wave then adds the probabilities in raster order from LDS.
#define SMF_WORDS 5  // staged words per window row: 4 + one for the byte alignment
__global__ __launch_bounds__(256) void k_update_smf(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.y]];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int b = blockIdx.x * 4 + wave;
  __shared__ uint32_t s_si[4][12 * SMF_WORDS];
  __shared__ __align__(16) float s_v[4][84];  // rows of 336 bytes: float4 reads stay aligned
  // updateSMF walks the SIR's WZ blocks; the barrier below is reached by all
  const bool on = b < f.N && wz_block(f, b);
  const int W = f.W;
  const SmfWin s = smf_window(f, on ? b : 0);
  const int cL = s.sL & ~3;
  if (on && lane < 12 * SMF_WORDS) {
    const int r = lane / SMF_WORDS, w = lane - r * SMF_WORDS, x = cL + 4 * w;
    s_si[wave][lane] = s.sT + r <= s.sB + 3 && x + 3 < W ? *(const uint32_t *)(f.si + (size_t)(s.sT + r) * W + x) : 0u;
  }
  uint32_t aw[4];
  uint32_t a2 = 0;
#pragma unroll
  for (int m = 0; m < 4; m++) {
    aw[m] = on ? *(const uint32_t *)(f.recons + (size_t)(s.top + m) * W + s.left) : 0u;
    a2 = __builtin_amdgcn_udot4(aw[m], aw[m], a2, false);
  }
  __syncthreads();
  float v[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int p = lane + 64 * h;
    v[h] = 0.0f;
    if (on && p < s.len) {
      const int pr = p / s.w, xo = s.sL + (p - pr * s.w) - cL, j = xo >> 2, sh = xo & 3;
      const uint32_t *R = s_si[wave] + pr * SMF_WORDS + j;
      uint32_t bb = 0, ab = 0;
#pragma unroll
      for (int ii = 0; ii < 4; ii++) {
        const uint32_t bw = __builtin_amdgcn_alignbyte(R[ii * SMF_WORDS + 1], R[ii * SMF_WORDS], sh);
        bb = __builtin_amdgcn_udot4(bw, bw, bb, false);
        ab = __builtin_amdgcn_udot4(aw[ii], bw, ab, false);
      }
      const float sse = (float)(a2 + bb - 2u * ab);
      v[h] = dvc_expf(-0.01f * sse);
      s_v[wave][p] = v[h];
    }
  }
  __syncthreads();
  if (!on) return;
  // the sum in raster order, the same on every lane (LDS broadcast reads)
  const int len = s.len;
  const float *V = s_v[wave];
  float sum = 0.0f;
  int p = 0;
  for (; p + 4 <= len; p += 4) {
    const float4 q = *(const float4 *)(V + p);
    sum += q.x;
    sum += q.y;
    sum += q.z;
    sum += q.w;
  }
  for (; p < len; p++) sum += V[p];
  float *pr = f.smf + (size_t)81 * b;
  const bool norm = (double)sum > 0.0;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int q = lane + 64 * h;
    if (q < s.len) pr[q] = norm ? v[h] / sum : (q == s.self ? 1.0f : v[h]);
  }
}

// update_si_ML (motionLearning.cpp:129-160): SMF-weighted over-complete
// coefficients for the bands after the frame's current band, accumulated in
// window raster order and truncated.  A wave owns ML_ROWS block rows of 16
// adjacent block columns; lane 4c + g owns row g (coefficients 4g..4g+3) of
// block column c and reads it from each window position's 64-byte record with
// one 16-byte load.  The wave streams down the over-complete rows its blocks'
// windows cover, reading each row once: the windows of block rows R-1, R and
// R+1 share position row 4R, so while position rows 4j..4j+3 stream by, block
// row j-1 takes its last row (at 4j), block row j its rows 5-8 and block row
// j+1 its rows 1-4, each in its own accumulators.  Every coefficient's sum
// still runs over its block's window in raster order; coefficients of bands
// already decoded are summed too but not stored.  A window row's weights are
// loaded three per lane by the four lanes of a block column and passed
// between them with quad permutes.
#define ML_ROWS 8
#define ML_COLS 16  // block columns per wave
template <int L>
__device__ __forceinline__ float quad_bcast(float v) {
  // quad_perm [L, L, L, L]: every lane of a quad takes the quad's lane L
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), L * 0x55, 0xf, 0xf, false));
}
__device__ __forceinline__ float quad_bcast(float v, int l) {  // l is a constant after unrolling
  switch (l) {
    case 0: return quad_bcast<0>(v);
    case 1: return quad_bcast<1>(v);
    case 2: return quad_bcast<2>(v);
    default: return quad_bcast<3>(v);
  }
}
__global__ __launch_bounds__(256) void k_update_si_ml(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.y]];
  const int bandStart = f.band + 1;  // update_si_ML(..., band+1, ...) (DVC_Decoder.cpp:369)
  if (bandStart >= 16) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int W = f.W, H = f.H, bw = W / 4, bh = H / 4, ow = W - 3;
  const int colTiles = (bw + ML_COLS - 1) / ML_COLS, rowTiles = (bh + ML_ROWS - 1) / ML_ROWS;
  const int tile = blockIdx.x * 4 + wave;
  if (tile >= colTiles * rowTiles) return;  // uniform over the wave
  const int ty = tile / colTiles, tx = tile - ty * colTiles;
  const int c = lane >> 2, g = lane & 3, bc = tx * ML_COLS + c;
  const bool live = bc < bw;
  int need = 0;  // coefficients 4g + jj still to predict
  for (int k = bandStart; k < 16; k++)
    if ((c_zigzag[k] >> 2) == g) need |= 1 << (c_zigzag[k] & 3);
  const int left = 4 * (live ? bc : 0), sL = imax(0, left - 4), w = imin(W - 4, left + 4) - sL + 1;
  const int R0 = ty * ML_ROWS, R1 = imin(bh, R0 + ML_ROWS);
  const int mS = imax(0, 4 * R0 - 4), mE = imin(H - 4, 4 * (R1 - 1) + 4);
  // one position row m for block row R (R0 <= R < R1): acc[jj] += p_q * o_q[jj]
  // over the row.  Positions past a narrow window's edge (q >= w) have p and o
  // zero: adding their +0 products leaves every sum's value unchanged (the
  // reference never adds them; a zero's sign cannot survive the truncation).
  auto rowTerms = [&](float *acc, int R, int m, const float (*o)[4]) {
    const float *pr = f.smf + (size_t)81 * (live ? R * bw + bc : 0) + (m - imax(0, 4 * R - 4)) * w;
    // lane g holds the weights q = g, g + 4, g + 8 of its block column's row
    float wv[3];
#pragma unroll
    for (int i = 0; i < 3; i++) wv[i] = live && g + 4 * i < w ? pr[g + 4 * i] : 0.0f;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      const float p = quad_bcast(wv[q >> 2], q & 3);
#pragma unroll
      for (int jj = 0; jj < 4; jj++) acc[jj] += p * o[q][jj];
    }
  };
  auto finish = [&](const float *acc, int R) {
    const int b = R * bw + bc;
    if (live && wz_block(f, b)) {
      int *dst = f.sideTrans + (4 * R + g) * W + left;
#pragma unroll
      for (int jj = 0; jj < 4; jj++)
        if (need >> jj & 1) dst[jj] = (int)acc[jj];
    }
  };
  float accP[4] = {}, accQ[4] = {};  // block rows j-1 and j, carried into chunk j
  for (int j = mS >> 2; 4 * j <= mE; j++) {
    float accN[4] = {};  // block row j+1, whose window starts at position row 4j
    for (int t = 0; t < 4; t++) {
      const int m = 4 * j + t;
      if (m < mS || m > mE) continue;  // uniform
      float o[9][4];
      const int4 *op = (const int4 *)(f.oct + (size_t)16 * ((size_t)m * ow + sL)) + g;
#pragma unroll
      for (int q = 0; q < 9; q++) {
        const int4 v = live && q < w ? op[4 * q] : make_int4(0, 0, 0, 0);
        o[q][0] = (float)v.x; o[q][1] = (float)v.y; o[q][2] = (float)v.z; o[q][3] = (float)v.w;
      }
      if (t == 0 && j - 1 >= R0) {  // the last row of block row j-1
        rowTerms(accP, j - 1, m, o);
        finish(accP, j - 1);
      }
      if (j >= R0 && j < R1) {
        rowTerms(accQ, j, m, o);
        if (m == H - 4) finish(accQ, j);  // the bottom block row's window ends at H-4
      }
      if (j + 1 < R1) rowTerms(accN, j + 1, m, o);
    }
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
      accP[jj] = accQ[jj];
      accQ[jj] = accN[jj];
    }
  }
}

// ------------------------------------------------------------ launchers
hipError_t launch_keyframe_filters(const uint8_t *in, uint8_t *lp, uint8_t *half, int W, int H, hipStream_t s) {
  hipLaunchKernelGGL(k_keyframe_filters, dim3((2 * W + 127) / 128, H), dim3(128), 0, s, in, lp, half, W, H);
  return hipGetLastError();
}
hipError_t launch_keyframe_batch(const KeyFrameJob *jobs, int n, int maxW, int maxH, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_keyframe_batch, dim3((2 * maxW + 127) / 128, maxH, n), dim3(128), 0, s, jobs);
  return hipGetLastError();
}
hipError_t launch_create_si(const FrameDev *F, const int *list, int n, int maxMB, CostTables ct, hipStream_t s) {
  hipLaunchKernelGGL(k_forward_me, dim3(maxMB, n), dim3(256), 0, s, F, list, ct);
  hipLaunchKernelGGL(k_si_chain, dim3(n), dim3(1024), 0, s, F, list, 0);
  return hipGetLastError();
}
hipError_t launch_forward_me(const FrameDev *F, const int *list, int n, int maxMB, CostTables ct, hipStream_t s) {
  hipLaunchKernelGGL(k_forward_me, dim3(maxMB, n), dim3(256), 0, s, F, list, ct);
  hipLaunchKernelGGL(k_si_chain, dim3(n), dim3(1024), 0, s, F, list, 1);
  return hipGetLastError();
}
hipError_t launch_noise_param(const FrameDev *F, const int *list, int n, int, hipStream_t s) {
  hipLaunchKernelGGL(k_noise_param, dim3(n), dim3(1024), 0, s, F, list);
  return hipGetLastError();
}
hipError_t launch_oct(const FrameDev *F, const int *list, int n, int maxPos, int initSideTrans, hipStream_t s) {
  hipLaunchKernelGGL(k_oct, dim3((maxPos + 255) / 256, n), dim3(256), 0, s, F, list, initSideTrans);
  return hipGetLastError();
}
// copySkipBlkToRecons (skipBlock.cpp:3-22): one thread per 8x8 block.
__global__ void k_skip_copy(const FrameDev *F, const int *list) {
  const FrameDev &f = F[list[blockIdx.y]];
  if (!f.mode) return;
  const int w8 = f.W >> 3, n8 = w8 * (f.H >> 3);
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n8 || f.mode[k] != kSkipBlock) return;
  const int top = (k / w8) * 8, left = (k % w8) * 8;
  for (int m = 0; m < 8; m++) {
    const size_t o = (size_t)(top + m) * f.W + left;
    for (int n = 0; n < 8; n++) f.recons[o + n] = f.past[o + n];  // caller buffers need not be aligned
  }
}
hipError_t launch_skip_copy(const FrameDev *F, const int *list, int n, int maxBlk88, hipStream_t s) {
  hipLaunchKernelGGL(k_skip_copy, dim3((maxBlk88 + 255) / 256, n), dim3(256), 0, s, F, list);
  return hipGetLastError();
}
hipError_t launch_oct_plane(const FrameDev *F, const int *list, int n, int maxPos, int index, hipStream_t s) {
  hipLaunchKernelGGL(k_oct_plane, dim3((maxPos + 255) / 256, n), dim3(256), 0, s, F, list, index);
  return hipGetLastError();
}
hipError_t launch_noise_dist(const FrameDev *F, const int *list, int n, int maxN, int maxStride, hipStream_t s) {
  (void)maxStride;  // one wave per block whatever the band's value count
  hipLaunchKernelGGL(k_noise_dist, dim3(maxN, n), dim3(ND_T), 0, s, F, list);
  return hipGetLastError();
}
hipError_t launch_condbit(const FrameDev *F, const int *list, int n, int maxN, hipStream_t s) {
  hipLaunchKernelGGL(k_condbit, dim3((maxN + 63) / 64, n), dim3(64), 0, s, F, list);
  return hipGetLastError();
}
hipError_t launch_put_bitplane(const FrameDev *F, const int *list, const BpFrameLink *links, int n, int maxN,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_put_bitplane, dim3((maxN + 255) / 256, n), dim3(256), 0, s, F, list, links);
  return hipGetLastError();
}
hipError_t launch_reconstruct(const FrameDev *F, const int *list, int n, int maxN, hipStream_t s) {
  hipLaunchKernelGGL(k_reconstruct, dim3((maxN + 255) / 256, n), dim3(256), 0, s, F, list);
  return hipGetLastError();
}
hipError_t launch_inverse_transform(const FrameDev *F, const int *list, int n, int maxN, hipStream_t s) {
  hipLaunchKernelGGL(k_inverse_transform, dim3((maxN + 255) / 256, n), dim3(256), 0, s, F, list);
  return hipGetLastError();
}
hipError_t launch_update_si(const FrameDev *F, const int *list, int n, int maxN, CostTables ct, hipStream_t s) {
  hipLaunchKernelGGL(k_update_si, dim3((maxN + 3) / 4, n), dim3(256), 0, s, F, list, ct);
  return hipGetLastError();
}
hipError_t launch_smf_init(const FrameDev *F, const int *list, int n, int maxN, hipStream_t s) {
  hipLaunchKernelGGL(k_smf_init, dim3((maxN + 255) / 256, n), dim3(256), 0, s, F, list);
  return hipGetLastError();
}
hipError_t launch_update_smf(const FrameDev *F, const int *list, int n, int maxN, hipStream_t s) {
  hipLaunchKernelGGL(k_update_smf, dim3((maxN + 3) / 4, n), dim3(256), 0, s, F, list);
  return hipGetLastError();
}
hipError_t launch_update_si_ml(const FrameDev *F, const int *list, int n, int maxBw, int maxBh, hipStream_t s) {
  const int tiles = (maxBw + ML_COLS - 1) / ML_COLS * ((maxBh + ML_ROWS - 1) / ML_ROWS);
  hipLaunchKernelGGL(k_update_si_ml, dim3((tiles + 3) / 4, n), dim3(256), 0, s, F, list);
  return hipGetLastError();
}

}  // namespace dvc
