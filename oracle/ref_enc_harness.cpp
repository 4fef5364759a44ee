// ref_enc_harness.cpp -- drives the reference encoder's own sources
// (DvcEnc/src/*.cpp, compiled from /root/reference by oracle/Makefile into
// oracle/_ref/libdvcencref.so) for one WZ frame: forwardTransform with the AC
// range, quantize, extractBitplane and LdpcaEncoder (DVC_Encoder.cpp:16-109
// with skip and intra blocks off).  TEST INFRASTRUCTURE ONLY.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "global.h"
#include "transform.h"
#include "quantization.h"
#include "bitplane.h"
#include "LdpcaEncoder.h"
#include "ReadLdpcaLadderFile_LdpcaEncoder.h"
#include "skipBlock.h"
#include "sJBIG.h"
#include "DVC_Encoder.h"
#include "frame.h"
#include "wzbitstream.h"
#include "intraBitstream.h"
#include "modeSelection.h"

extern "C" {
int refenc_setup(int W, int H, int qindex, const char *ladder) {
  memset(&codingOption, 0, sizeof(codingOption));
  codingOption.FrameWidth = W;
  codingOption.FrameHeight = H;
  codingOption.Qindex = (unsigned char)qindex;
  FrameSize = W * H;
  Block44Num = FrameSize / 16;
  Block88Num = FrameSize / 64;
  Block44Height = H / 4; Block44Width = W / 4;
  Block88Height = H / 8; Block88Width = W / 8;
  Qtable = QuantizationTable[qindex];
  QtableBits = QuantizationTableBits[qindex];
  QbitsNumPerBlk = 0;
  for (int j = 0; j < 16; j++) QbitsNumPerBlk += QtableBits[j];
  for (int j = 0; j < 16; j++) ACLevelShift[j] = (j > 0 && Qtable[j] > 0) ? (Qtable[j] - 2) / 2 : 0;
  initTransform();
  ReadLadder_encoder((char *)ladder);
  return QbitsNumPerBlk;
}

// Encodes one frame: acRange[16], crc[nbp], acc[nbp * N/8].  Returns nbp.
int refenc_encode(const unsigned char *frame, int *acRange, unsigned char *crc, unsigned char *acc) {
  EncodedFrame e;
  memset(&e, 0, sizeof(e));
  e.skipBlockFlag = (unsigned char *)calloc(Block88Num, 1);
  e.intraBlockFlag = (unsigned char *)calloc(Block88Num, 1);
  int *trans = (int *)malloc(sizeof(int) * FrameSize);
  forwardTransform((unsigned char *)frame, trans, &e);
  quantize(trans, &e);
  unsigned char **bp = (unsigned char **)malloc(sizeof(unsigned char *) * QbitsNumPerBlk);
  for (int i = 0; i < QbitsNumPerBlk; i++) bp[i] = (unsigned char *)malloc(Block44Num);
  extractBitplane(trans, bp, &e);
  int nb = Block44Num / 8;
  for (int i = 0; i < QbitsNumPerBlk; i++) {
    int len = 0;
    unsigned char *a = LdpcaEncoder(bp[i], &len, &crc[i]);
    memcpy(acc + (size_t)i * nb, a, nb);
    free(a);
    free(bp[i]);
  }
  memcpy(acRange, e.ACRange, sizeof(int) * 16);
  free(bp); free(trans); free(e.skipBlockFlag); free(e.intraBlockFlag);
  return QbitsNumPerBlk;
}

// The same with skip blocks (DVC_Encoder.cpp:24-28 with EnableSkipBlock):
// findSkipBlock against the past key frame `ref`, then the frame's fields as
// above; flags[Block88Num] receives skipBlockFlag, and when the frame keeps
// skip mode its map's JBIG data as the encoder writes it to the bitstream
// (wzbitstream.cpp:36-38) goes to jbig (cap bytes), its length to *jbig_len.
// Returns skipMode.
int refenc_encode_skip(const unsigned char *frame, const unsigned char *ref, int *acRange, unsigned char *crc,
                       unsigned char *acc, unsigned char *flags, unsigned char *jbig, int cap, int *jbig_len) {
  EncodedFrame e;
  memset(&e, 0, sizeof(e));
  e.skipBlockFlag = (unsigned char *)calloc(Block88Num, 1);
  e.intraBlockFlag = (unsigned char *)calloc(Block88Num, 1);
  findSkipBlock((unsigned char *)frame, (unsigned char *)ref, &e);
  int *trans = (int *)malloc(sizeof(int) * FrameSize);
  forwardTransform((unsigned char *)frame, trans, &e);
  quantize(trans, &e);
  unsigned char **bp = (unsigned char **)malloc(sizeof(unsigned char *) * QbitsNumPerBlk);
  for (int i = 0; i < QbitsNumPerBlk; i++) bp[i] = (unsigned char *)malloc(Block44Num);
  extractBitplane(trans, bp, &e);
  int nb = Block44Num / 8;
  for (int i = 0; i < QbitsNumPerBlk; i++) {
    int len = 0;
    unsigned char *a = LdpcaEncoder(bp[i], &len, &crc[i]);
    memcpy(acc + (size_t)i * nb, a, nb);
    free(a);
    free(bp[i]);
  }
  memcpy(acRange, e.ACRange, sizeof(int) * 16);
  memcpy(flags, e.skipBlockFlag, Block88Num);
  *jbig_len = 0;
  if (e.skipMode == 1) {
    int nbytes = 0;
    unsigned char *bits = BitToBitmap(e.skipBlockFlag, Block88Width, Block88Height, &nbytes);
    FILE *fp = tmpfile();
    unsigned char jbgh[20];
    const int n = simple_JBIG_enc(fp, bits, Block88Width, Block88Height, jbgh, Block88Height, 1);
    rewind(fp);
    *jbig_len = n <= cap ? (int)fread(jbig, 1, n, fp) : -1;
    fclose(fp);
    free(bits);
  }
  const int mode = e.skipMode;
  free(bp); free(trans); free(e.skipBlockFlag); free(e.intraBlockFlag);
  return mode;
}

// The reference encoder's main loop (DvcEnc main.cpp:20-100) on frames in
// memory: every GOP through encodeGOP with the given options, the key frames'
// reconstructions being the originals; writes the WZ bitstream to wz_path and,
// with intra mode on, the intra bitstream to intra_path.  Returns the number
// of WZ frames.
int refenc_run(const unsigned char *frames, int nframes, int W, int H, int q, int gop, int skip, int intra,
               int deblock, int intra_qp, const char *ladder, const char *wz_path, const char *intra_path) {
  refenc_setup(W, H, q, ladder);  // geometry, tables, ladder, transform
  codingOption.FramesToBeEncoded = nframes;
  codingOption.GOP = (unsigned char)gop;
  codingOption.EnableSkipBlock = skip;
  codingOption.EnableIntraMode = intra;
  codingOption.IntraQP = (unsigned char)intra_qp;
  codingOption.DeblockingMode = (unsigned char)deblock;
  Block44Num = W * H / 16; Block88Num = W * H / 64;
  initTransform();  // again, now that the intra options are set (it builds the intra tables only then)
  initFrameBuffer();
  for (int i = 0; i < nframes; i++) {
    memcpy(oriFrames[i], frames + (size_t)i * FrameSize, FrameSize);
    memcpy(reconsFrames[i], frames + (size_t)i * FrameSize, FrameSize);
  }
  if (intra) {
    IntraFp = fopen(intra_path, "wb+");
    if (!IntraFp) return -1;
    intrap.thita = 0.0;
    intrap.kappa = 0.0;
    init_IntraBuffer();
  }
  int wzCount = 0;
  for (int i = 0; i < nframes; i += gop) {
    int end = i + gop;
    if (end >= nframes) end = nframes - 1;
    encodeGOP(i, end, &wzCount);
  }
  writeWZbitstream((char *)wz_path);
  if (intra) free_IntraBuffer();  // flushes the last partial byte and closes IntraFp
  freeFrameBuffer();
  Kmatrix = NULL;  // freeFrameBuffer frees it, but only intra-mode runs allocate it again
  return wzCount;
}
}
