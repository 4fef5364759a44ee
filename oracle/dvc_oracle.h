/* dvc_oracle.h -- CPU restatement of the reference decoder's Wyner-Ziv hot
 * path (jcluox/dvc, DvcDec ParallelMode 0 and 1) and of the WZ part of its
 * encoder.  TEST INFRASTRUCTURE ONLY: it is the checker that the HIP path is
 * compared against (tests/, the smoke check and bench.py's CPU baseline leg);
 * the product never links or calls it.
 *
 * Every function names the reference lines it restates.  Floating-point
 * expressions keep the reference's operand order and precision (float where
 * the C++ overloads resolve to float, double where the reference uses double)
 * and call the host libm, so on the same C library the results are the
 * reference's bit for bit.  Build flags: -O2 -ffp-contract=off (oracle/Makefile).
 *
 * Parity pinning (how each part is checked, see tests/):
 *   ladder/inverse-H   : derived QCIF ladder and inverse equal the reference's
 *                        RegDecoder1584.lad / Inverse_Matrix_H_Reg1584.bin (hash
 *                        fixtures); inv(H)*H = I over GF(2) for both sizes.
 *   CRC-8              : check value CRC("123456789") = 0xF4, table = poly 0x07.
 *   syndrome/encoder   : H*x = s on the accepted level; prefix-XOR identity.
 *   BP, CE, rate search: against the reference's own LdpcaDecoder.cpp compiled
 *                        from source (oracle/_ref) on seeded inputs, and golden
 *                        per-bitplane levels from full reference decodes.
 *   SI, CNM, SMF, ML,
 *   transform, recon   : against the reference sources compiled into
 *                        oracle/_ref on seeded frames, brute-force ME on tiny
 *                        frames, and golden hashes/PSNR from full decodes.
 */
#ifndef DVC_ORACLE_H
#define DVC_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ LDPCA */
/* Decoder ladder: ReadLadder_decoder (ReadLdpcaLadderFile_LdpcaDecoder.cpp:131-250). */
typedef struct {
  int numCodes, N, E, period;
  int *jc;            /* [N+1] */
  int *bitsPerPeriod; /* [numCodes] */
  int *idx;           /* [numCodes*period] transmission pattern per code */
  int *ir;            /* [numCodes*E] decoder row of every edge, per code */
  int *syndLength;    /* [numCodes] m = (N/period)*b */
  float *codeRate;    /* [numCodes] (float)m/(float)N */
  uint8_t *invH;      /* [N * 4*ceil(N/32)] MSB-first packed rows of H^-1 */
  int invWords;       /* ceil(N/32) */
} orc_ladder;

/* Encoder ladder: ReadLadder_encoder (DvcEnc/src/ReadLdpcaLadderFile_LdpcaEncoder.cpp:12-36). */
typedef struct {
  int N, E;
  int *jc; /* [N+1] */
  int *ir; /* [E] encoder (unmerged) rows */
} orc_enc_ladder;

int orc_read_decoder_ladder(orc_ladder *L, const char *lad_path, const char *inv_path);
int orc_read_encoder_ladder(orc_enc_ladder *L, const char *lad_path);
/* Regenerate a decoder ladder from an encoder ladder and the per-code
 * transmission pattern (Appendix-A derivation: ir_dec = rowmap(ir_enc)). */
int orc_derive_decoder_ladder(orc_ladder *L, const orc_enc_ladder *enc, int numCodes, int period,
                              const int *bitsPerPeriod, const int *idx);
/* GF(2) inverse of the encoder parity matrix, packed like the reference's
 * Inverse_Matrix_H_Reg*.bin. Returns 0 on success, -1 if singular. */
int orc_invert_h(const orc_enc_ladder *enc, uint8_t *out, int words_per_row);
void orc_free_ladder(orc_ladder *L);
void orc_free_enc_ladder(orc_enc_ladder *L);

/* BitToByte / ByteToBit (LdpcaDecoder.cpp:41-107): MSB-first packing. */
int orc_bits_to_bytes(const uint8_t *bits, int nbits, uint8_t *bytes);
void orc_bytes_to_bits(const uint8_t *bytes, int nbits, uint8_t *bits);
/* CRC8 (crc8.h:20-67): poly x^8+x^2+x+1, init 0, no reflection, xorout 0. */
uint8_t orc_crc8(const uint8_t *buf, int len);

/* ComputeCE (LdpcaDecoder.cpp:127-143): sequential float sum. */
float orc_compute_ce(const float *LR, int N);
/* ComputeCE_OpenMP (LdpcaDecoder.cpp:109-126) with nthreads (0 = default). */
float orc_compute_ce_omp(const float *LR, int N, int nthreads);

/* Syndrome formation (LdpcaDecoder.cpp:684-693) from N accumulated-syndrome bits. */
void orc_form_syndrome(const orc_ladder *L, int code, const uint8_t *acc_bits, uint8_t *synd);

/* beliefPropagation (LdpcaDecoder.cpp:462-580): mode 0.  Returns 1 when the
 * syndrome check passes.  *iters = iterations run (1-based count). */
int orc_bp_mode0(const orc_ladder *L, int code, const float *LLR, const uint8_t *synd,
                 uint8_t *decoded, int max_iter, int *iters);
/* beliefPropagation_OpenMP (LdpcaDecoder.cpp:164-306): mode 1, with the
 * negative early jump that compares against the carried-over `decoded`. */
int orc_bp_mode1(const orc_ladder *L, int code, const float *LLR, const uint8_t *synd,
                 uint8_t *decoded, int max_iter, int nthreads, int *iters);
/* invH_mul (LdpcaDecoder.cpp:624-643). */
void orc_invh_mul(const orc_ladder *L, const uint8_t *synd, uint8_t *decoded);

typedef struct {
  int start_level;   /* nCodeStart */
  int level;         /* accepted code index */
  int rate;          /* m + 8 */
  int requests;      /* codes tried (level - start_level + 1 with rate_step 0) */
  int total_iters;   /* BP iterations over all tried levels */
  int used_inverse;  /* 1 when the full-rate inverse produced the result */
} orc_ldpca_stats;

/* LdpcaDecoder (LdpcaDecoder.cpp:646-771).  LR is converted to LLR in place
 * exactly as the reference does (:663-667).  mode 0 or 1; `decoded` is in/out
 * (mode 1 compares against its previous content). */
void orc_ldpca_decode(const orc_ladder *L, float *LR, const uint8_t *acc_bits, uint8_t crc,
                      float entropy, int mode, int max_iter, int nthreads, uint8_t *decoded,
                      orc_ldpca_stats *st);
/* The same with rate_step 1: the rate search of a build with DYNAMICSTEP
 * defined (LdpcaDecoder.cpp:19-21, 669-672, 760-768). */
void orc_ldpca_decode_steps(const orc_ladder *L, float *LR, const uint8_t *acc_bits, uint8_t crc,
                            float entropy, int mode, int max_iter, int nthreads, int rate_step,
                            uint8_t *decoded, orc_ldpca_stats *st);

/* LdpcaEncoder (DvcEnc/src/LdpcaEncoder.cpp:64-95): accumulated syndrome
 * bytes (N/8) and CRC of the source bitplane. */
void orc_ldpca_encode(const orc_enc_ladder *L, const uint8_t *src_bits, uint8_t *acc_bytes,
                      uint8_t *crc);

/* --------------------------------------------------- frame-level geometry */
typedef struct {
  int W, H, frameSize;
  int b44n, b44w, b44h;
  int b88n, b88w, b88h;
  int b1616n, b1616w, b1616h;
} orc_geom;
void orc_geom_init(orc_geom *g, int W, int H);

/* Quantiser tables (reconstruction.cpp:9-78), zig-zag (global.cpp:36-41). */
extern const int orc_qtable[9][16];
extern const int orc_qtable_bits[9][16];
extern const int orc_zigzag[16];

typedef struct { double x, y; } orc_point;
typedef struct { orc_point past, future, vector; } orc_mv;

typedef struct {
  int searchTop, searchBottom, searchLeft, searchRight;
  int top, left, searchWidth, searchRange, blkIdx;
} orc_search_info;

typedef struct {
  float prob[81];
  int searchTop, searchLeft, searchRight, searchBottom;
  int searchWidth, blkIdx, probLen;
} orc_smf;

/* ----------------------------------------------- side-information creation */
void orc_lowpass(const uint8_t *in, int W, int H, uint8_t *out);            /* filter.cpp:214-233 */
void orc_fir(const uint8_t *in, int W, int H, uint8_t *out);                /* filter.cpp:55-117 */
void orc_init_1616(const orc_geom *g, orc_search_info *si1616);             /* sideInfoCreation.cpp:24-59 */
/* forwardME (sideInfoCreation.cpp:61-145): mcroBlockMV[b1616n] and holes. */
void orc_forward_me(const orc_geom *g, const orc_search_info *si1616, const uint8_t *fut_lp,
                    const uint8_t *past_lp, int past, int cur, int fut, orc_mv *mbmv,
                    orc_point *holes, int omp);
void orc_bidirect_me(const orc_geom *g, const uint8_t *fut_half, const uint8_t *past_half,
                     int past, int cur, int fut, orc_mv *mv, int blockLen);   /* :235-310 */
void orc_motion_filter_comp(const orc_geom *g, const uint8_t *fut_half, const uint8_t *past_half,
                            int past, int cur, int fut, orc_mv *blockMV, double *mse,
                            uint8_t *si, int *residue);                          /* :312-436 */
/* createSideInfo (:438-551) given the low-passed and interpolated key frames. */
void orc_create_side_info(const orc_geom *g, const uint8_t *past_lp, const uint8_t *fut_lp,
                          const uint8_t *past_half, const uint8_t *fut_half, int past, int cur,
                          int fut, uint8_t *si, int *residue, orc_mv *mbmv, orc_mv *blockMV,
                          int omp);
int orc_init_sir(const orc_geom *g, const uint8_t *blockModeFlag, orc_search_info *sir); /* :554-596 */
/* updateSI + subPelRefine (:598-736); mv_out (optional) gets pBestX,pBestY,fBestX,fBestY per block. */
void orc_update_si(const orc_geom *g, const uint8_t *recons, const uint8_t *past_full,
                   const uint8_t *fut_full, const uint8_t *past_half, const uint8_t *fut_half,
                   int past, int cur, int fut, const orc_search_info *sir, int nsir, uint8_t *si,
                   int *mv_out, int omp);

/* ----------------------------------------------------- transform and CNM */
void orc_forward44(int *blk);                        /* transform.cpp:63-104 */
void orc_inverse44(int *blk);                        /* transform.cpp:106-151 */
void orc_forward_transform(const orc_geom *g, const int *pix, int *trans); /* :153-190 */
void orc_inverse_transform(const orc_geom *g, const int *trans, uint8_t *recons,
                           const orc_search_info *sir, int nsir);         /* :234-273 */
void orc_overcomplete(const orc_geom *g, const uint8_t *pix, int *oct /* [(H-3)*(W-3)*16] */, int omp);
void orc_noise_parameter(const orc_geom *g, int qindex, int *trans, float *alpha,
                         const uint8_t *blockModeFlag);                  /* noiseModel.cpp:9-75 */
void orc_init_smf(const orc_geom *g, orc_smf *smf, const orc_search_info *sir, int nsir); /* motionLearning.cpp:6-47 */
void orc_init_si_ml(const orc_geom *g, int *si_ml, const int *oct);      /* :49-68 */
void orc_update_smf(const orc_geom *g, orc_smf *smf, const uint8_t *recons, const uint8_t *si,
                    const orc_search_info *sir, int nsir, int omp);     /* :70-127 */
void orc_update_si_ml(const orc_geom *g, int *sideTrans, const int *oct, int band,
                      const orc_smf *smf, const orc_search_info *sir, int nsir, int omp); /* :129-160 */
/* noiseDistribution (noiseModel.cpp:77-132): returns number of floats per block. */
int orc_noise_distribution(const orc_geom *g, int range, int indexInBlk, const float *alpha,
                           const orc_smf *smf, const int *oct, const orc_search_info *sir, int nsir,
                           float *dist, int omp);
void orc_cond_bit_prob(const orc_geom *g, int qindex, float *LR, int indexInBlk, int bit,
                       int stepSize, const int *reconsTrans, int range, const float *dist,
                       const orc_search_info *sir, int nsir, int omp);  /* noiseModel.cpp:191-310 */
void orc_put_bitplane(const orc_geom *g, const uint8_t *bitplane, int *reconsTrans, int indexInBlk,
                      const uint8_t *blockModeFlag);                    /* bitplane.cpp:4-22 */
void orc_reconstruction(const orc_geom *g, int qindex, const int *reconsTrans, int *sideTrans,
                        int indexInBlk, const float *alpha, const int *stepSize,
                        const uint8_t *blockModeFlag);                  /* reconstruction.cpp:184-289 */
double orc_psnr(const uint8_t *a, const uint8_t *b, int n);            /* frame.cpp:115-124 */

/* ------------------------------------------------------- WZ encoder part */
/* forwardTransform with AC range (DvcEnc/src/transform.cpp:229-276),
 * quantize (quantization.cpp:95-140), extractBitplane (bitplane.cpp:4-38),
 * LdpcaEncoder per bitplane.  Outputs ACRange[16], crc[nbp], acc[nbp][N/8]. */
int orc_encode_wz_frame(const orc_geom *g, int qindex, const orc_enc_ladder *enc,
                        const uint8_t *frame, int *acRange, uint8_t *crc, uint8_t *acc);
int orc_bitplanes_per_frame(int qindex);

/* ------------------------------------------------------- whole WZ frame */
typedef struct {
  int mode;        /* 0 sequential, 1 OpenMP */
  int nthreads;    /* OpenMP threads for mode 1 */
  int max_iter;    /* BP iteration cap (reference: 50) */
  int rate_step;   /* 1: the DYNAMICSTEP build's rate search (orc_ldpca_decode_steps) */
} orc_decode_opts;

typedef struct {
  int nbitplanes;
  int bitrate;                 /* ACRange bits + sum of (m+8) (+ 2 mode bits) */
  int levels[128];             /* accepted code index per bitplane */
  int start_levels[128];       /* nCodeStart per bitplane */
  int iters[128];              /* BP iterations per bitplane (all levels tried) */
  float entropy[128];
  double psnr;
  double t_si, t_oct, t_cnm, t_condbit, t_ldpca, t_updsi, t_ml, t_total; /* ms */
} orc_frame_result;

/* decodeGOP body for one WZ frame (DVC_Decoder.cpp:26-474) with skip, intra
 * and deblocking off.  `recons` receives the decoded frame; `orig` is used for
 * PSNR only. */
int orc_decode_wz_frame(const orc_geom *g, int qindex, const orc_ladder *L, const uint8_t *past,
                        const uint8_t *fut, int past_idx, int cur_idx, int fut_idx,
                        const int *acRange, const uint8_t *crc, const uint8_t *acc,
                        const uint8_t *orig, const orc_decode_opts *opt, uint8_t *recons,
                        orc_frame_result *res);

#ifdef __cplusplus
}
#endif
#endif
