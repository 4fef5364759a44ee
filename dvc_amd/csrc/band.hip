// band.hip -- the per-band entry points of include/dvc_accel.h: the calls a
// decoder that keeps its own per-frame loop makes per band, on the
// reference's data layouts (SMF, SearchInfo, Trans, MV of global.h), run by
// the same kernels the batched frame decoder launches (frame_kernels.hip) on
// a one-frame descriptor held in the context.
#include <string.h>
#include <algorithm>
#include "dvc_internal.h"
#include "frame_kernels.h"

namespace dvc {

// One WZ frame's device state for the per-band calls, and a second set of
// buffers for dvc_forward_me (which runs before the frame's set-up).
struct BandState {
  int W = 0, H = 0, N = 0, qindex = 0;
  bool ready = false;
  uint8_t *mem = nullptr;  // one block: frame buffers, descriptor and list
  size_t bytes = 0;
  float *dist = nullptr;   // the band's noise distribution, N * stride floats
  size_t distBytes = 0;
  int distIndex = -1, distRange = -1;  // band of the distribution in `dist`
  dvc_smf *smfRef = nullptr;           // reference-layout SMF staging (one per SIR entry)
  int nsir = 0;                        // SIR entries (WZ 4x4 blocks)
  int *d_blk = nullptr;                // blkIdx of each SIR entry (device)
  uint8_t *gat = nullptr;              // the distribution in noiseDistribution's SIR-entry layout
  size_t gatBytes = 0;
  FrameDev fd{};
  FrameDev *d_fd = nullptr;
  int *d_list = nullptr;
  uint8_t *fmeMem = nullptr;  // dvc_forward_me's block
  size_t fmeBytes = 0;
};

void band_free(dvc_ctx *c) {
  if (!c->band) return;
  if (c->band->mem) (void)hipFree(c->band->mem);
  if (c->band->dist) (void)hipFree(c->band->dist);
  if (c->band->fmeMem) (void)hipFree(c->band->fmeMem);
  if (c->band->gat) (void)hipFree(c->band->gat);
  delete c->band;
  c->band = nullptr;
}

static size_t al(size_t v) { return (v + 255) & ~(size_t)255; }

// Device block of at least `bytes` in *p (grown, contents dropped).
static int grow(uint8_t **p, size_t *have, size_t bytes) {
  if (*have >= bytes) return DVC_OK;
  if (*p) DVC_HIP(hipFree(*p));
  *p = nullptr;
  *have = 0;
  DVC_HIP(hipMalloc((void **)p, bytes));
  *have = bytes;
  return DVC_OK;
}

static bool valid_size(int W, int H) { return (W == 176 && H == 144) || (W == 352 && H == 288); }

// prob[0..probLen) of the reference-layout SMFs (one per SIR entry i) to the
// kernels' N x 81 layout (block blk[i]) and back.
__global__ void k_smf_unpack(const dvc_smf *src, float *dst, int n, const int *blk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int len = min(src[i].probLen, 81);
  float *d = dst + (size_t)81 * blk[i];
  for (int p = 0; p < 81; p++) d[p] = p < len ? src[i].prob[p] : 0.0f;
}
__global__ void k_smf_pack(const float *src, dvc_smf *dst, int n, const int *blk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int len = min(dst[i].probLen, 81);
  const float *s = src + (size_t)81 * blk[i];
  for (int p = 0; p < len; p++) dst[i].prob[p] = s[p];
}
// The block-major distribution to noiseDistribution's layout: SIR entry i's
// `stride` values at i * stride.
__global__ void k_dist_gather(const float *src, float *dst, int n, const int *blk, int stride) {
  const int i = blockIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || k >= stride) return;
  dst[(size_t)i * stride + k] = src[(size_t)blk[i] * stride + k];
}

static int upload_fd(dvc_ctx *c, BandState *B) {
  DVC_HIP(hipMemcpyAsync(B->d_fd, &B->fd, sizeof(FrameDev), hipMemcpyHostToDevice, c->stream));
  return DVC_OK;
}

static int need_frame(dvc_ctx *c, const char *who) {
  DVC_CHECK(c, DVC_EINVAL, std::string(who) + ": NULL context");
  DVC_CHECK(c->band && c->band->ready, DVC_ESTATE, std::string(who) + ": call dvc_band_frame_init first");
  DVC_HIP(hipSetDevice(c->device));
  return DVC_OK;
}

}  // namespace dvc

using namespace dvc;

extern "C" int dvc_band_frame_init(dvc_ctx *c, int W, int H, int qindex, const uint8_t *past, const uint8_t *fut,
                                   int past_idx, int cur_idx, int fut_idx, const dvc_search_info *sir, int nsir) {
  DVC_CHECK(c, DVC_EINVAL, "dvc_band_frame_init: NULL context");
  DVC_CHECK(valid_size(W, H), DVC_EINVAL, "dvc_band_frame_init: size must be QCIF or CIF");
  DVC_CHECK(qindex >= 1 && qindex <= 8, DVC_EINVAL, "dvc_band_frame_init: qindex must be 1..8");
  DVC_CHECK(past && fut, DVC_EINVAL, "dvc_band_frame_init: NULL key frame");
  DVC_CHECK(past_idx < cur_idx && cur_idx < fut_idx, DVC_EINVAL, "dvc_band_frame_init: need past_idx < cur_idx < fut_idx");
  const int N = W * H / 16, bw = W / 4, n8 = N / 4;
  // initSIR (sideInfoCreation.cpp:554-596) lists the 4x4 blocks of the WZ 8x8
  // blocks in raster order; the blocks it leaves out (skip and intra) are
  // marked non-WZ in an 8x8 mode map, which every per-block kernel tests
  std::vector<int> blk;
  std::vector<uint8_t> mode;
  if (sir) {
    DVC_CHECK(nsir >= 0 && nsir <= N, DVC_EINVAL, "dvc_band_frame_init: more SIR entries than blocks");
    mode.assign(n8, 2);
    std::vector<int> per8(n8, 0);
    for (int i = 0; i < nsir; i++) {
      const int b = sir[i].blkIdx;
      DVC_CHECK(b >= 0 && b < N && (i == 0 || b > sir[i - 1].blkIdx) && sir[i].top == (b / bw) * 4 &&
                    sir[i].left == (b % bw) * 4,
                DVC_EINVAL, "dvc_band_frame_init: the SIR must list 4x4 blocks in raster order (blkIdx, top, left)");
      const int b8 = (b / bw / 2) * (bw / 2) + (b % bw) / 2;
      mode[b8] = 0;
      per8[b8]++;
      blk.push_back(b);
    }
    for (int k = 0; k < n8; k++)
      DVC_CHECK(per8[k] == 0 || per8[k] == 4, DVC_EINVAL,
                "dvc_band_frame_init: the SIR lists part of an 8x8 block (block modes are per 8x8 block)");
  } else {
    nsir = N;
  }
  if (blk.empty() && nsir == N)
    for (int b = 0; b < N; b++) blk.push_back(b);
  if (nsir == N) mode.clear();  // every block is WZ
  DVC_HIP(hipSetDevice(c->device));
  int rc = ensure_cost_tables(c);
  if (rc) return rc;
  if (!c->band) c->band = new BandState();
  BandState *B = c->band;
  B->ready = false;
  const size_t FS = (size_t)W * H;
  size_t off = 0;
  auto take = [&off](size_t b) { const size_t o = off; off = al(off + b); return o; };
  const size_t oPast = take(FS), oFut = take(FS), oPlp = take(FS), oFlp = take(FS), oPh = take(4 * FS),
               oFh = take(4 * FS), oSi = take(FS), oRec = take(FS), oRt = take(4 * FS), oAl = take(4 * FS),
               oOct = take((size_t)64 * (W - 3) * (H - 3)),
               oOb = take((size_t)4 * (W - 3) * (H - 3)), oSmf = take((size_t)4 * 81 * N),
               oSmfRef = take(sizeof(dvc_smf) * (size_t)N), oLR = take(4 * (size_t)N), oFd = take(sizeof(FrameDev)),
               oList = take(sizeof(int)), oBlk = take(sizeof(int) * (size_t)N), oMode = take((size_t)n8);
  if ((rc = grow(&B->mem, &B->bytes, off))) return rc;
  uint8_t *m = B->mem;
  FrameDev &f = B->fd;
  memset(&f, 0, sizeof(f));
  f.W = W; f.H = H; f.N = N; f.qindex = qindex;
  f.iPC = (double)(cur_idx - past_idx) / (double)(fut_idx - past_idx);
  f.iFC = (double)(fut_idx - cur_idx) / (double)(fut_idx - past_idx);
  f.past = m + oPast; f.fut = m + oFut; f.past_lp = m + oPlp; f.fut_lp = m + oFlp;
  f.past_half = m + oPh; f.fut_half = m + oFh;
  f.si = m + oSi; f.recons = m + oRec; f.reconsTrans = (int *)(m + oRt); f.alpha = (float *)(m + oAl);
  f.oct = (int *)(m + oOct); f.octBand = (int *)(m + oOb); f.smf = (float *)(m + oSmf); f.LR = (float *)(m + oLR);
  B->smfRef = (dvc_smf *)(m + oSmfRef);
  B->d_blk = (int *)(m + oBlk);
  B->nsir = nsir;
  f.mode = mode.empty() ? nullptr : m + oMode;
  B->d_fd = (FrameDev *)(m + oFd);
  B->d_list = (int *)(m + oList);
  B->W = W; B->H = H; B->N = N; B->qindex = qindex;
  B->distIndex = B->distRange = -1;
  hipStream_t s = c->stream;
  const int zero = 0;
  DVC_HIP(hipMemcpyAsync(B->d_list, &zero, sizeof(int), hipMemcpyHostToDevice, s));
  if (nsir > 0) DVC_HIP(hipMemcpyAsync(B->d_blk, blk.data(), sizeof(int) * (size_t)nsir, hipMemcpyHostToDevice, s));
  if (!mode.empty()) DVC_HIP(hipMemcpyAsync((void *)f.mode, mode.data(), (size_t)n8, hipMemcpyHostToDevice, s));
  DVC_HIP(hipMemcpyAsync((void *)f.past, past, FS, hipMemcpyHostToDevice, s));
  DVC_HIP(hipMemcpyAsync((void *)f.fut, fut, FS, hipMemcpyHostToDevice, s));
  DVC_HIP(launch_keyframe_filters(f.past, (uint8_t *)f.past_lp, (uint8_t *)f.past_half, W, H, s));
  DVC_HIP(launch_keyframe_filters(f.fut, (uint8_t *)f.fut_lp, (uint8_t *)f.fut_half, W, H, s));
  if ((rc = upload_fd(c, B))) return rc;
  DVC_HIP(hipStreamSynchronize(s));
  B->ready = true;
  return DVC_OK;
}

extern "C" int dvc_noise_distribution(dvc_ctx *c, int range, int index, const float *alpha, const dvc_smf *smf,
                                      const int *oct, float *dist) {
  int rc = need_frame(c, "dvc_noise_distribution");
  if (rc) return rc;
  DVC_CHECK(index >= 0 && index < 16 && range > 0 && alpha && smf && oct, DVC_EINVAL,
            "dvc_noise_distribution: bad argument");
  BandState *B = c->band;
  FrameDev &f = B->fd;
  const int W = B->W, H = B->H, N = B->N, stride = index ? 2 * range + 1 : range;
  hipStream_t s = c->stream;
  const size_t need = sizeof(float) * (size_t)N * stride;
  B->distIndex = B->distRange = -1;
  if ((rc = grow((uint8_t **)&B->dist, &B->distBytes, need))) return rc;
  DVC_HIP(hipMemcpyAsync(f.alpha, alpha, sizeof(float) * W * H, hipMemcpyHostToDevice, s));
  DVC_HIP(hipMemcpyAsync(f.oct, oct, sizeof(int) * 16 * (size_t)(W - 3) * (H - 3), hipMemcpyHostToDevice, s));
  const int ns = B->nsir;
  if (ns > 0) {
    DVC_HIP(hipMemcpyAsync(B->smfRef, smf, sizeof(dvc_smf) * (size_t)ns, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_smf_unpack, dim3((ns + 255) / 256), dim3(256), 0, s, B->smfRef, f.smf, ns, B->d_blk);
    DVC_HIP(hipGetLastError());
  }
  f.index = index; f.range = range; f.distStride = stride; f.dist = B->dist;
  if ((rc = upload_fd(c, B))) return rc;
  DVC_HIP(launch_oct_plane(B->d_fd, B->d_list, 1, (W - 3) * (H - 3), index, s));
  DVC_HIP(launch_noise_dist(B->d_fd, B->d_list, 1, N, stride, s));
  if (dist && ns == N) {
    DVC_HIP(hipMemcpyAsync(dist, B->dist, need, hipMemcpyDeviceToHost, s));
  } else if (dist && ns > 0) {  // SIR entries only, in their order
    const size_t out = sizeof(float) * (size_t)ns * stride;
    if ((rc = grow(&B->gat, &B->gatBytes, out))) return rc;
    hipLaunchKernelGGL(k_dist_gather, dim3((stride + 255) / 256, ns), dim3(256), 0, s, B->dist, (float *)B->gat, ns,
                       B->d_blk, stride);
    DVC_HIP(hipGetLastError());
    DVC_HIP(hipMemcpyAsync(dist, B->gat, out, hipMemcpyDeviceToHost, s));
  }
  DVC_HIP(hipStreamSynchronize(s));
  B->distIndex = index;
  B->distRange = range;
  return DVC_OK;
}

extern "C" int dvc_cond_bit_prob(dvc_ctx *c, float *LR, int index, int bit, int step, const int *recons_trans,
                                 int range) {
  int rc = need_frame(c, "dvc_cond_bit_prob");
  if (rc) return rc;
  DVC_CHECK(LR && recons_trans && bit >= 1 && bit <= 30 && step > 0, DVC_EINVAL, "dvc_cond_bit_prob: bad argument");
  BandState *B = c->band;
  DVC_CHECK(B->distIndex == index && B->distRange == range, DVC_ESTATE,
            "dvc_cond_bit_prob: no noise distribution of this band (call dvc_noise_distribution first)");
  FrameDev &f = B->fd;
  hipStream_t s = c->stream;
  DVC_HIP(hipMemcpyAsync(f.reconsTrans, recons_trans, sizeof(int) * B->W * B->H, hipMemcpyHostToDevice, s));
  f.index = index; f.bit = bit; f.step = step; f.range = range;
  f.levelShift = index ? (qtable_entry(B->qindex, index) - 2) / 2 : 0;
  f.distStride = index ? 2 * range + 1 : range;
  f.dist = B->dist;
  if ((rc = upload_fd(c, B))) return rc;
  DVC_HIP(launch_condbit(B->d_fd, B->d_list, 1, B->N, s));
  DVC_HIP(hipMemcpyAsync(LR, f.LR, sizeof(float) * B->N, hipMemcpyDeviceToHost, s));
  DVC_HIP(hipStreamSynchronize(s));
  return DVC_OK;
}

extern "C" int dvc_update_si(dvc_ctx *c, uint8_t *side_info, const uint8_t *recons) {
  int rc = need_frame(c, "dvc_update_si");
  if (rc) return rc;
  DVC_CHECK(side_info && recons, DVC_EINVAL, "dvc_update_si: NULL argument");
  BandState *B = c->band;
  FrameDev &f = B->fd;
  hipStream_t s = c->stream;
  const size_t FS = (size_t)B->W * B->H;
  DVC_HIP(hipMemcpyAsync(f.si, side_info, FS, hipMemcpyHostToDevice, s));
  DVC_HIP(hipMemcpyAsync(f.recons, recons, FS, hipMemcpyHostToDevice, s));
  DVC_HIP(launch_update_si(B->d_fd, B->d_list, 1, B->N, CostTables{c->d_costFme, c->d_costUsi}, s));
  DVC_HIP(hipMemcpyAsync(side_info, f.si, FS, hipMemcpyDeviceToHost, s));
  DVC_HIP(hipStreamSynchronize(s));
  return DVC_OK;
}

extern "C" int dvc_update_smf(dvc_ctx *c, dvc_smf *smf, const uint8_t *recons, const uint8_t *side_info) {
  int rc = need_frame(c, "dvc_update_smf");
  if (rc) return rc;
  DVC_CHECK(smf && recons && side_info, DVC_EINVAL, "dvc_update_smf: NULL argument");
  BandState *B = c->band;
  FrameDev &f = B->fd;
  hipStream_t s = c->stream;
  const size_t FS = (size_t)B->W * B->H;
  const int N = B->N, ns = B->nsir;
  if (ns == 0) return DVC_OK;  // no WZ block: updateSMF's loop is empty
  DVC_HIP(hipMemcpyAsync(B->smfRef, smf, sizeof(dvc_smf) * (size_t)ns, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_smf_unpack, dim3((ns + 255) / 256), dim3(256), 0, s, B->smfRef, f.smf, ns, B->d_blk);
  DVC_HIP(hipGetLastError());
  DVC_HIP(hipMemcpyAsync(f.si, side_info, FS, hipMemcpyHostToDevice, s));
  DVC_HIP(hipMemcpyAsync(f.recons, recons, FS, hipMemcpyHostToDevice, s));
  DVC_HIP(launch_update_smf(B->d_fd, B->d_list, 1, N, s));
  hipLaunchKernelGGL(k_smf_pack, dim3((ns + 255) / 256), dim3(256), 0, s, f.smf, B->smfRef, ns, B->d_blk);
  DVC_HIP(hipGetLastError());
  DVC_HIP(hipMemcpyAsync(smf, B->smfRef, sizeof(dvc_smf) * (size_t)ns, hipMemcpyDeviceToHost, s));
  DVC_HIP(hipStreamSynchronize(s));
  return DVC_OK;
}

extern "C" int dvc_forward_me(dvc_ctx *c, int W, int H, const uint8_t *fut_lp, const uint8_t *past_lp, int past_idx,
                              int cur_idx, int fut_idx, dvc_mv *mb_mv) {
  DVC_CHECK(c, DVC_EINVAL, "dvc_forward_me: NULL context");
  DVC_CHECK(valid_size(W, H) && fut_lp && past_lp && mb_mv, DVC_EINVAL, "dvc_forward_me: bad argument");
  DVC_CHECK(past_idx < cur_idx && cur_idx < fut_idx, DVC_EINVAL, "dvc_forward_me: need past_idx < cur_idx < fut_idx");
  static_assert(sizeof(dvc_mv) == sizeof(MVd), "MV layouts differ");
  DVC_HIP(hipSetDevice(c->device));
  int rc = ensure_cost_tables(c);
  if (rc) return rc;
  if (!c->band) c->band = new BandState();
  BandState *B = c->band;
  const size_t FS = (size_t)W * H, nmb = FS / 256;
  size_t off = 0;
  auto take = [&off](size_t b) { const size_t o = off; off = al(off + b); return o; };
  const size_t oPlp = take(FS), oFlp = take(FS), oFwd = take(sizeof(MVd) * nmb), oMb = take(sizeof(MVd) * nmb),
               oHoles = take(16 * nmb), oFd = take(sizeof(FrameDev)), oList = take(sizeof(int));
  if ((rc = grow(&B->fmeMem, &B->fmeBytes, off))) return rc;
  uint8_t *m = B->fmeMem;
  FrameDev f;
  memset(&f, 0, sizeof(f));
  f.W = W; f.H = H; f.N = (int)(FS / 16);
  f.iPC = (double)(cur_idx - past_idx) / (double)(fut_idx - past_idx);
  f.iFC = (double)(fut_idx - cur_idx) / (double)(fut_idx - past_idx);
  f.past_lp = m + oPlp; f.fut_lp = m + oFlp;
  f.fwd = (MVd *)(m + oFwd); f.mbmv = (MVd *)(m + oMb); f.holes = (double *)(m + oHoles);
  hipStream_t s = c->stream;
  const int zero = 0;
  DVC_HIP(hipMemcpyAsync(m + oList, &zero, sizeof(int), hipMemcpyHostToDevice, s));
  DVC_HIP(hipMemcpyAsync(m + oFd, &f, sizeof(FrameDev), hipMemcpyHostToDevice, s));
  DVC_HIP(hipMemcpyAsync(m + oPlp, past_lp, FS, hipMemcpyHostToDevice, s));
  DVC_HIP(hipMemcpyAsync(m + oFlp, fut_lp, FS, hipMemcpyHostToDevice, s));
  DVC_HIP(launch_forward_me((const FrameDev *)(m + oFd), (const int *)(m + oList), 1, (int)nmb,
                            CostTables{c->d_costFme, c->d_costUsi}, s));
  DVC_HIP(hipMemcpyAsync(mb_mv, f.mbmv, sizeof(MVd) * nmb, hipMemcpyDeviceToHost, s));
  DVC_HIP(hipStreamSynchronize(s));
  return DVC_OK;
}
