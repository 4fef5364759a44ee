// ref_accel_stubs.cpp -- the reference decoder (DvcDec, its own unmodified
// objects) with its hot-path calls routed to the MI355X library through the
// C ABI of include/dvc_accel.h, by link-time wrapping (-Wl,--wrap) instead of
// a source edit.  This is the binding a DvcDec maintainer adds; oracle/Makefile
// links it with the reference's objects and oracle/ref_harness.cpp into
// oracle/_ref/libdvcref_accel.so, and a GPU test runs the reference's own
// decodeGOP through it against the golden frames.  TEST INFRASTRUCTURE ONLY.
//
// Routed calls (callers in DVC_Decoder.cpp, defined in other objects):
//   LdpcaDecoder       (DVC_Decoder.cpp:267)  -> dvc_ldpca_decode
//   initSMF            (:86)                  -> (reference) + dvc_band_frame_init
//                      (the place of CNM_CUDA_Init_Variable/ME_CUDA_Init_Variable, :91-95)
//   noiseDistribution  (:165)                 -> dvc_noise_distribution
//   condBitProb        (:195)                 -> dvc_cond_bit_prob
//   updateSI           (:323)                 -> dvc_update_si
//   updateSMF          (:373)                 -> dvc_update_smf
// createSideInfo is wrapped only to learn the WZ frame's key-frame numbers.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "global.h"
#include "ReadLdpcaLadderFile_LdpcaDecoder.h"
#include "../include/dvc_accel.h"

extern int LDPC_iterations;

static dvc_ctx *g_ctx = NULL;
static int g_past = 0, g_cur = 1, g_fut = 2;
static long g_routed[6];  // calls routed: ldpca, frame init, noise dist, condbit, updateSI, updateSMF

static char g_err[512];  // first failure of a routed call ("" while none)

// Records the first failure; the caller checks ref_accel_error() after the
// decode (a wrapped call has no way to return an error to its caller).
static void fail(const char *where) {
  if (!g_err[0]) snprintf(g_err, sizeof(g_err), "%s: %s", where, dvc_last_error());
}

extern "C" {
// Creates the library context (InitCUDA + LDPCA_CUDA_Init_Variable, main.cpp:92-104).
int ref_accel_open(int max_iter) {
  if (g_ctx) return 0;
  dvc_params p;
  memset(&p, 0, sizeof(p));
  p.max_iter = max_iter;
  p.parallel_mode = 0;
  memset(g_routed, 0, sizeof(g_routed));
  g_err[0] = 0;
  return dvc_ctx_create(&p, &g_ctx);
}
int ref_accel_close(void) {
  int rc = dvc_ctx_destroy(g_ctx);
  g_ctx = NULL;
  return rc;
}
long ref_accel_routed(int k) { return k >= 0 && k < 6 ? g_routed[k] : -1; }
const char *ref_accel_error(void) { return g_err; }

// LdpcaDecoder(LR_intrinsic, accumulatedSyndrome (one byte per bit, ByteToBit's
// output), decoded, rate, crc, RateLowerBound) (LdpcaDecoder.cpp:646-771).
void dvc_stub_LdpcaDecoder(float *LR, unsigned char *accBits, unsigned char *decoded, int *rate, unsigned char crc,
                           float entropy) {
  const int N = codeLength;
  std::vector<uint8_t> acc(N / 8, 0);
  for (int k = 0; k < N; k++) acc[k >> 3] |= (uint8_t)((accBits[k] & 1) << (7 - (k & 7)));
  int start = 0;  // nCodeStart from the caller's entropy (:657-659)
  while (codeRate[start] < entropy) start++;
  dvc_bp_stats st;
  if (dvc_ldpca_decode(g_ctx, N, LR, acc.data(), crc, start, decoded, &st) != DVC_OK) {
    fail("dvc_ldpca_decode");
    *rate = 0;
    return;
  }
  *rate = st.rate;
  LDPC_iterations += st.iterations;
  g_routed[0]++;
}

void __real__Z14createSideInfoiiiP8SideInfo(int past, int future, int current, SideInfo *si);
void __wrap__Z14createSideInfoiiiP8SideInfo(int past, int future, int current, SideInfo *si) {
  g_past = past;
  g_fut = future;
  g_cur = current;
  __real__Z14createSideInfoiiiP8SideInfo(past, future, current, si);
}

void __real__Z7initSMFP3SMFP3SIR(SMF *smf, SIR *sir);
void __wrap__Z7initSMFP3SMFP3SIR(SMF *smf, SIR *sir) {
  __real__Z7initSMFP3SMFP3SIR(smf, sir);
  if (dvc_band_frame_init(g_ctx, codingOption.FrameWidth, codingOption.FrameHeight, codingOption.Qindex,
                          decodedFrames[g_past].reconsFrame, decodedFrames[g_fut].reconsFrame, g_past, g_cur, g_fut,
                          (const dvc_search_info *)sir->searchInfo, sir->searchCandidate) != DVC_OK)
    fail("dvc_band_frame_init");
  g_routed[1]++;
}

// The distribution stays in the library context, as d_distribution stays on
// the device in the CUDA build; the caller only frees the returned pointer.
float *__wrap__Z17noiseDistributioniiPfP3SMFP5TransP3SIR(int range, int indexInBlk, float *alpha, SMF *smf,
                                                          Trans *oct, SIR *sir) {
  (void)sir;
  if (dvc_noise_distribution(g_ctx, range, indexInBlk, alpha, (const dvc_smf *)smf, (const int *)oct, NULL) != DVC_OK)
    fail("dvc_noise_distribution");
  g_routed[2]++;
  return (float *)malloc(sizeof(float));
}

void __wrap__Z11condBitProbPfiiiPiiS_P3SIR(float *LR, int indexInBlk, int bit, int stepSize, int *reconsTrans,
                                           int range, float *noiseDist, SIR *sir) {
  (void)noiseDist;
  (void)sir;
  if (dvc_cond_bit_prob(g_ctx, LR, indexInBlk, bit, stepSize, reconsTrans, range) != DVC_OK) fail("dvc_cond_bit_prob");
  g_routed[3]++;
}

void __wrap__Z8updateSIP8SideInfoPhiiiP3SIR(SideInfo *si, unsigned char *recons, int current, int past, int future,
                                            SIR *sir) {
  (void)current; (void)past; (void)future; (void)sir;
  if (dvc_update_si(g_ctx, si->sideInfoFrame, recons) != DVC_OK) fail("dvc_update_si");
  g_routed[4]++;
}

void __wrap__Z9updateSMFP3SMFPhS1_P3SIR(SMF *smf, unsigned char *recons, unsigned char *sideInfo, SIR *sir) {
  (void)sir;
  if (dvc_update_smf(g_ctx, (dvc_smf *)smf, recons, sideInfo) != DVC_OK) fail("dvc_update_smf");
  g_routed[5]++;
}
}  // extern "C"
