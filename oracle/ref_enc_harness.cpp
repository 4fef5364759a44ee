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
}
