// ref_harness.cpp -- a thin driver around the reference decoder's own CPU
// sources (DvcDec/src/*.cpp, compiled from /root/reference by oracle/Makefile
// into oracle/_ref/libdvcref.so).  It sets the reference's global state for
// one frame geometry and exposes its functions with plain-pointer C entry
// points, so tests can check the oracle's restatement against the reference
// on seeded inputs.  TEST INFRASTRUCTURE ONLY; built only where the reference
// tree exists.
//
// The reference's LDPCA decoder is reached through -Wl,--wrap=LdpcaDecoder and
// -Wl,--wrap=ComputeCE(_OpenMP) so that decodeGOP's per-bitplane results (rate,
// entropy) can be recorded without editing the reference.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "global.h"
#include "LdpcaDecoder.h"
#include "ReadLdpcaLadderFile_LdpcaDecoder.h"
#include "sideInfoCreation.h"
#include "filter.h"
#include "noiseModel.h"
#include "motionLearning.h"
#include "transform.h"
#include "reconstruction.h"
#include "bitplane.h"
#include "frame.h"
#include "DVC_Decoder.h"
#include "jbig.h"
#include "sJBIG.h"
#include "wzbitstream.h"
#include "intraBitstream.h"
#include <vector>

// Iteration counter that the reference's beliefPropagation_OpenMP increments
// (declared extern at LdpcaDecoder.cpp:30 and frame.cpp:10); the harness owns
// it and reports it through ref_bp_iterations().
int LDPC_iterations = 0;

static int g_nframes = 0;
static unsigned char *g_flags = NULL;  // all-WZ block-mode map
static int g_log_rate[4096], g_log_level[4096], g_log_n = 0;
static float g_log_entropy[4096];
static int g_ent_n = 0;

// The wrapped functions are C++ (mangled) symbols; the linker's --wrap works on
// those names, so the wrapper and the real entry carry the mangled suffixes.
#ifdef DVC_ACCEL
// the accelerated build: LdpcaDecoder goes to the GPU library (ref_accel_stubs.cpp)
#define REAL_LDPCA dvc_stub_LdpcaDecoder
#else
#define REAL_LDPCA __real__Z12LdpcaDecoderPfPhS0_Pihf
#endif
#define REAL_CE __real__Z9ComputeCEPf
#define REAL_CE_OMP __real__Z16ComputeCE_OpenMPPf
extern "C" {
void REAL_LDPCA(float *, unsigned char *, unsigned char *, int *, unsigned char, float);
float REAL_CE(float *);
float REAL_CE_OMP(float *);

void __wrap__Z12LdpcaDecoderPfPhS0_Pihf(float *LR, unsigned char *acc, unsigned char *decoded, int *rate,
                         unsigned char crc, float entropy) {
  REAL_LDPCA(LR, acc, decoded, rate, crc, entropy);
  if (g_log_n < 4096) {
    g_log_rate[g_log_n] = *rate;
    int lvl = -1;
    for (int c = 0; c < numCodes; c++)
      if (syndLength[c] + 8 == *rate) { lvl = c; break; }
    g_log_level[g_log_n] = lvl;
    g_log_n++;
  }
}
float __wrap__Z9ComputeCEPf(float *p) {
  float e = REAL_CE(p);
  if (g_ent_n < 4096) g_log_entropy[g_ent_n++] = e;
  return e;
}
float __wrap__Z16ComputeCE_OpenMPPf(float *p) {
  float e = REAL_CE_OMP(p);
  if (g_ent_n < 4096) g_log_entropy[g_ent_n++] = e;
  return e;
}

int ref_setup(int W, int H, int qindex, int mode, int nthreads, int nframes) {
  memset(&codingOption, 0, sizeof(codingOption));
  codingOption.FrameWidth = W;
  codingOption.FrameHeight = H;
  codingOption.FramesToBeDecoded = nframes;
  codingOption.ParallelMode = mode;
  codingOption.UseCoreNum = nthreads;
  codingOption.Qindex = (unsigned char)qindex;
  codingOption.GOP = 2;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  FrameSize = W * H;
  Block44Num = FrameSize / 16;
  Block88Num = Block44Num / 4;
  Block1616Num = Block88Num / 4;
  Block44Height = H / 4; Block44Width = W / 4;
  Block88Height = H / 8; Block88Width = W / 8;
  Block1616Height = H / 16; Block1616Width = W / 16;
  Qtable = QuantizationTable[qindex];
  QtableBits = QuantizationTableBits[qindex];
  QbitsNumPerBlk = 0;
  for (int j = 0; j < 16; j++) QbitsNumPerBlk += QtableBits[j];
  for (int j = 0; j < 16; j++) ACLevelShift[j] = (j > 0 && Qtable[j] > 0) ? (Qtable[j] - 2) / 2 : 0;
  g_nframes = nframes;
  initFrameBuffer();
  init_1616ME(SearchInfo1616);
  initTransform();
  encodedFrames = (EncodedFrame *)calloc(nframes, sizeof(EncodedFrame));
  for (int i = 0; i < nframes; i++) {
    decodedFrames[i].hasReferenced = 0;
    decodedFrames[i].isWZFrame = 0;
    decodedFrames[i].bitRate = 0;
  }
  free(g_flags);
  g_flags = (unsigned char *)calloc(Block88Num, 1);
  return 0;
}

int ref_load_ladder(const char *lad, const char *inv) {
  ReadLadder_decoder((char *)lad, (char *)inv);
  initLdpcaBuffer();
  return 0;
}

int ref_num_codes(void) { return numCodes; }
int ref_synd_length(int c) { return syndLength[c]; }
int ref_ir(int c, int e) { return ir[c][e]; }
int ref_bp_iterations(void) { return LDPC_iterations; }

void ref_set_frame(int idx, const unsigned char *pix, const unsigned char *orig) {
  memcpy(decodedFrames[idx].reconsFrame, pix, FrameSize);
  if (orig) memcpy(oriFrames[idx], orig, FrameSize);
  decodedFrames[idx].hasReferenced = 0;
}

void ref_set_encoded(int idx, const int *acRange, const unsigned char *crc, const unsigned char *acc) {
  EncodedFrame *e = &encodedFrames[idx];
  memcpy(e->ACRange, acRange, sizeof(int) * 16);
  e->BlockModeFlag = (unsigned char *)calloc(Block88Num, 1);
  e->skipMode = 0; e->intraMode = 0; e->skipBlockBits = 1; e->intraBlockBits = 1;
  e->CRC = (unsigned char *)malloc(QbitsNumPerBlk);
  memcpy(e->CRC, crc, QbitsNumPerBlk);
  e->AccumulatedSyndrome = (unsigned char **)malloc(sizeof(unsigned char *) * QbitsNumPerBlk);
  int nb = Block44Num / 8;
  for (int j = 0; j < QbitsNumPerBlk; j++) {
    e->AccumulatedSyndrome[j] = (unsigned char *)malloc(nb);
    memcpy(e->AccumulatedSyndrome[j], acc + (size_t)j * nb, nb);
  }
}

// decodeGOP (DVC_Decoder.cpp:26-474) for one (begin, end) pair.
int ref_decode_gop(int begin, int end) {
  g_log_n = 0; g_ent_n = 0;
  decodeGOP(begin, end);
  return g_log_n;
}
void ref_get_frame(int idx, unsigned char *out) { memcpy(out, decodedFrames[idx].reconsFrame, FrameSize); }
double ref_get_psnr(int idx) { return decodedFrames[idx].PSNR; }
int ref_get_bitrate(int idx) { return decodedFrames[idx].bitRate; }
int ref_get_log(int *rates, int *levels, float *entropy) {
  memcpy(rates, g_log_rate, sizeof(int) * g_log_n);
  memcpy(levels, g_log_level, sizeof(int) * g_log_n);
  memcpy(entropy, g_log_entropy, sizeof(float) * g_ent_n);
  return g_log_n;
}

// ---- function-level entry points
int ref_bp(int mode, int code, float *LLR, unsigned char *synd, unsigned char *decoded) {
  int m = syndLength[code];
  return mode == 0 ? beliefPropagation(LLR, synd, decoded, m, code)
                   : beliefPropagation_OpenMP(LLR, synd, decoded, m, code);
}
void ref_ldpca(float *LR, unsigned char *acc_bits, unsigned char crc, float entropy,
               unsigned char *decoded, int *rate) {
  REAL_LDPCA(LR, acc_bits, decoded, rate, crc, entropy);
}
float ref_compute_ce(float *LR) { return REAL_CE(LR); }
float ref_compute_ce_omp(float *LR) { return REAL_CE_OMP(LR); }

void ref_lowpass(const unsigned char *in, unsigned char *out) {
  lowPassFilter((unsigned char *)in, codingOption.FrameWidth, codingOption.FrameHeight, out);
}
void ref_fir(const unsigned char *in, unsigned char *out) {
  FIR_filter((unsigned char *)in, codingOption.FrameWidth, codingOption.FrameHeight, out);
}

// createSideInfo (sideInfoCreation.cpp:438-551): key frames must be set with
// ref_set_frame.  mbmv / bmv receive 6 doubles per MV (past, future, vector).
void ref_create_side_info(int past, int fut, int cur, unsigned char *si, int *residue, double *mbmv,
                          double *bmv) {
  createSideInfo(past, fut, cur, &sideInfo);
  memcpy(si, sideInfo.sideInfoFrame, FrameSize);
  memcpy(residue, sideInfo.residueFrame, sizeof(int) * FrameSize);
  if (mbmv) memcpy(mbmv, sideInfo.mcroBlockMV, sizeof(MV) * Block1616Num);
  if (bmv) memcpy(bmv, sideInfo.BlockMV, sizeof(MV) * Block88Num);
}

static SIR make_sir(void) {
  SIR sir;
  sir.searchInfo = (SearchInfo *)malloc(sizeof(SearchInfo) * Block44Num);
  initSIR(&sir, g_flags);
  return sir;
}

// updateSI (sideInfoCreation.cpp:642-736) after ref_create_side_info set the
// interpolated key frames.  si is read and written.
void ref_update_si(int cur, int past, int fut, const unsigned char *recons, unsigned char *si, int omp) {
  SIR sir = make_sir();
  memcpy(sideInfo.sideInfoFrame, si, FrameSize);
  if (omp) updateSI_OpenMP(&sideInfo, (unsigned char *)recons, cur, past, fut, &sir);
  else updateSI(&sideInfo, (unsigned char *)recons, cur, past, fut, &sir);
  memcpy(si, sideInfo.sideInfoFrame, FrameSize);
  free(sir.searchInfo);
}

void ref_forward_transform(const int *pix, int *out) { forwardTransform((int *)pix, out); }
void ref_inverse_transform(const int *trans, unsigned char *recons) {
  SIR sir = make_sir();
  inverseTransform((int *)trans, recons, &sir);
  free(sir.searchInfo);
}
void ref_noise_parameter(int *trans, float *alpha) { noiseParameter(trans, alpha, g_flags); }
void ref_overcomplete(const unsigned char *pix, int *oct, int omp) {
  if (omp) overCompleteTransform_OpenMP((unsigned char *)pix, (Trans *)oct, codingOption.FrameHeight - 4, codingOption.FrameWidth - 4);
  else overCompleteTransform((unsigned char *)pix, (Trans *)oct, codingOption.FrameHeight - 4, codingOption.FrameWidth - 4);
}
int ref_init_smf(void *smf) {
  SIR sir = make_sir();
  initSMF((SMF *)smf, &sir);
  int n = sir.searchCandidate;
  free(sir.searchInfo);
  return n;
}
void ref_init_si_ml(int *si_ml, const int *oct) { init_si_ML(si_ml, (Trans *)oct); }
void ref_update_smf(void *smf, const unsigned char *recons, const unsigned char *si, int omp) {
  SIR sir = make_sir();
  if (omp) updateSMF_OpenMP((SMF *)smf, (unsigned char *)recons, (unsigned char *)si, &sir);
  else updateSMF((SMF *)smf, (unsigned char *)recons, (unsigned char *)si, &sir);
  free(sir.searchInfo);
}
void ref_update_si_ml(int *sideTrans, const int *oct, int band, void *smf, int omp) {
  SIR sir = make_sir();
  if (omp) update_si_ML_OpenMP(sideTrans, (Trans *)oct, band, (SMF *)smf, &sir);
  else update_si_ML(sideTrans, (Trans *)oct, band, (SMF *)smf, &sir);
  free(sir.searchInfo);
}
int ref_noise_distribution(int range, int index, float *alpha, void *smf, const int *oct, float *out, int omp) {
  SIR sir = make_sir();
  float *d = omp ? noiseDistribution_OpenMP(range, index, alpha, (SMF *)smf, (Trans *)oct, &sir)
                 : noiseDistribution(range, index, alpha, (SMF *)smf, (Trans *)oct, &sir);
  int per = index == 0 ? range : 2 * range + 1;
  memcpy(out, d, sizeof(float) * (size_t)per * sir.searchCandidate);
  free(d);
  free(sir.searchInfo);
  return per;
}
void ref_cond_bit_prob(float *LR, int index, int bit, int step, int *reconsTrans, int range, float *dist, int omp) {
  SIR sir = make_sir();
  if (omp) condBitProb_OpenMP(LR, index, bit, step, reconsTrans, range, dist, &sir);
  else condBitProb(LR, index, bit, step, reconsTrans, range, dist, &sir);
  free(sir.searchInfo);
}
void ref_put_bitplane(const unsigned char *bp, int *reconsTrans, int index) {
  putBitplaneToCoeff((unsigned char *)bp, reconsTrans, index, g_flags);
}
void ref_reconstruction(int *reconsTrans, int *sideTrans, int index, float *alpha, int *stepSize) {
  unsigned char *db = (unsigned char *)malloc(FrameSize);
  reconstruction(reconsTrans, sideTrans, index, alpha, stepSize, g_flags, db);
  free(db);
}
double ref_psnr(const unsigned char *a, const unsigned char *b) { return psnr((unsigned char *)a, (unsigned char *)b); }

// ---- JBIG block-mode maps and the WZ bitstream reader (fixture generation)
static std::vector<unsigned char> g_bie;
static void harness_bie_out(unsigned char *start, size_t len, void *) { g_bie.insert(g_bie.end(), start, start + len); }

// The reference's JBIG-KIT encoder on a 0/1 map (rows packed MSB first):
// the 20-byte header followed by the data into out; returns the length, or
// -1 when cap is too small.
int ref_jbig_encode(const unsigned char *map, int w, int h, int order, int options, int l0, int mx,
                    unsigned char *out, int cap) {
  const int bpl = (w + 7) / 8;
  std::vector<unsigned char> packed((size_t)bpl * h, 0);
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++)
      if (map[y * w + x]) packed[(size_t)y * bpl + x / 8] |= (unsigned char)(0x80 >> (x & 7));
  unsigned char *planes[1] = {packed.data()};
  struct jbg_enc_state se;
  g_bie.clear();
  jbg_enc_init(&se, w, h, 1, planes, harness_bie_out, NULL);
  jbg_enc_options(&se, order, options, l0, mx, 0);
  unsigned char hdr[20];
  jbg_enc_out(&se, hdr);
  jbg_enc_free(&se);
  if ((int)(20 + g_bie.size()) > cap) return -1;
  memcpy(out, hdr, 20);
  memcpy(out + 20, g_bie.data(), g_bie.size());
  return (int)(20 + g_bie.size());
}

// The reference's JBIG-KIT decoder on a whole BIE: image (w*h bytes 0/1)
// and the bytes consumed; returns 0, or the JBIG-KIT error code.
int ref_jbig_decode_bie(const unsigned char *bie, int len, unsigned char *img, int cap, int *w, int *h, int *consumed) {
  struct jbg_dec_state sd;
  jbg_dec_init(&sd);
  size_t cnt = 0;
  int r = jbg_dec_in(&sd, (unsigned char *)bie, len, &cnt);
  if (r != JBG_EOK && r != JBG_EOK_INTR) { jbg_dec_free(&sd); return r; }
  *w = (int)jbg_dec_getwidth(&sd);
  *h = (int)jbg_dec_getheight(&sd);
  *consumed = (int)cnt;
  const unsigned char *p = jbg_dec_getimage(&sd, 0);
  const int bpl = (*w + 7) / 8;
  if (*w * *h <= cap)
    for (int y = 0; y < *h; y++)
      for (int x = 0; x < *w; x++) img[y * *w + x] = (p[(size_t)y * bpl + x / 8] >> (7 - (x & 7))) & 1;
  jbg_dec_free(&sd);
  return 0;
}

// One block-mode map as readGOP decodes it (wzbitstream.cpp:40-66): the
// rebuilt header, simple_JBIG_dec over `bufsize` bytes, ByteToBitmap.
// Returns the decoded byte count (decode_byte).
int ref_jbig_map_decode(const unsigned char *bid, int len, int w, int h, unsigned char *map) {
  const int bufsize = ((((w + 4) >> 3) << 3) * h >> 3) * 2;
  struct jbg_enc_state se;
  jbg_enc_init(&se, w, h, 1, NULL, NULL, NULL);
  jbg_enc_options(&se, JBG_ILEAVE | JBG_SMID, JBG_TPDON | JBG_DPON, h, 8, 0);
  unsigned char jbgh[20];
  jbg_set_header(&se, jbgh);
  std::vector<unsigned char> enc((size_t)bufsize, 0), dec((size_t)bufsize, 0);
  memcpy(enc.data(), bid, len < bufsize ? len : bufsize);
  const int used = simple_JBIG_dec(enc.data(), bufsize, jbgh, bufsize, dec.data());
  unsigned char *bits = ByteToBitmap(dec.data(), w, h);
  memcpy(map, bits, (size_t)w * h);
  free(bits);
  return used;
}

// A frame's skip map as readGOP sets it from the map's JBIG data: skip
// blocks in BlockModeFlag, skipMode 1, skipBlockBits = decoded bytes * 8 + 1.
int ref_set_skip_map(int idx, const unsigned char *bid, int len) {
  EncodedFrame *e = &encodedFrames[idx];
  std::vector<unsigned char> map((size_t)Block88Num);
  const int used = ref_jbig_map_decode(bid, len, Block88Width, Block88Height, map.data());
  for (int i = 0; i < Block88Num; i++)
    if (map[i] == 1) e->BlockModeFlag[i] = SKIP_B;
  e->skipMode = 1;
  e->skipBlockBits = (used << 3) + 1;
  return used;
}

// readWZBitstream on a file: the header fields and, per WZ frame index,
// ACRange, the modes and mode bits, BlockModeFlag, CRCs and syndromes
// (fetched with ref_wz_frame).
int ref_read_wz_bitstream(const char *path, int *hdr) {
  memset(&codingOption, 0, sizeof(codingOption));  // its file-name fields are appended to
  QbitsNumPerBlk = 0;                              // readWZBitstream accumulates into it
  readWZBitstream((char *)path);
  hdr[0] = codingOption.FrameWidth; hdr[1] = codingOption.FrameHeight; hdr[2] = codingOption.Qindex;
  hdr[3] = codingOption.IntraQP; hdr[4] = codingOption.GOP; hdr[5] = codingOption.FramesToBeDecoded;
  hdr[6] = codingOption.DeblockingMode; hdr[7] = QbitsNumPerBlk;
  return 0;
}
// A whole encoded sequence as the reference's main.cpp loads it
// (main.cpp:36-42): the WZ bitstream, then the intra bitstream when a frame
// uses intra blocks.  Keeps the ParallelMode and thread count ref_setup set.
// Returns IntraModeExist.
int ref_load_stream(const char *wz_path, const char *intra_path, int *hdr) {
  const int mode = codingOption.ParallelMode, threads = codingOption.UseCoreNum;
  ref_read_wz_bitstream(wz_path, hdr);
  codingOption.ParallelMode = mode;
  codingOption.UseCoreNum = threads;
  if (codingOption.IntraModeExist) readIntraBitstream((char *)intra_path);
  return codingOption.IntraModeExist;
}
int ref_wz_frame(int idx, int *acRange, int *modes, unsigned char *flags, unsigned char *crc, unsigned char *acc) {
  EncodedFrame *e = &encodedFrames[idx];  // a WZ frame's index (key frames' entries are not set)
  memcpy(acRange, e->ACRange, sizeof(int) * 16);
  modes[0] = e->skipMode; modes[1] = e->intraMode; modes[2] = e->skipBlockBits; modes[3] = e->intraBlockBits;
  memcpy(flags, e->BlockModeFlag, Block88Num);
  memcpy(crc, e->CRC, QbitsNumPerBlk);
  const int nb = Block44Num / 8;
  for (int j = 0; j < QbitsNumPerBlk; j++) memcpy(acc + (size_t)j * nb, e->AccumulatedSyndrome[j], nb);
  return 0;
}
}  // extern "C"
