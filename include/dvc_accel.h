/* dvc_accel.h -- C ABI of the MI355X (gfx950) Wyner-Ziv decoding hot path.
 *
 * The reference decoder (jcluox/dvc, DvcDec) reaches its accelerator through a
 * set of extern "C" functions over file-scope global state, selected by
 * codingOption.ParallelMode 2/3:
 *   LDPCA   DvcDec/inc/LDPCA_cuda.h:16-21      InitCUDA, LDPCA_CUDA_Init_Variable,
 *                                              beliefPropagation_CUDA(_earlyJump)
 *   CNM     DvcDec/inc/noiseModel_kernel.h:13-19 CNM_CUDA_Init_Buffer, noiseDistribution_CUDA,
 *                                              condBitProb_CUDA
 *   SI      DvcDec/inc/updateSI_kernel.h:13-17, forwardME_kernel.h:11-13
 *                                              updateSI_CUDA, forwardME_CUDA
 *   SMF     DvcDec/inc/motionLearning_kernel.h:11-13 updateSMF_CUDA
 * This header replaces that set with the same operations around an opaque
 * context, so several frames and streams can be decoded at once.  All
 * functions return 0 on success and a negative DVC_E* code on failure (never
 * exit()); dvc_last_error() describes the last failure of the calling thread.
 * Host pointers are caller-owned and only read or written during the call.
 * A context is used by one host thread at a time.
 */
#ifndef DVC_ACCEL_H
#define DVC_ACCEL_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define DVC_OK 0
#define DVC_EINVAL (-1)   /* bad argument (size, code length, pointer) */
#define DVC_EDEVICE (-2)  /* HIP runtime error or no gfx950 device */
#define DVC_EIO (-3)      /* ladder pack or cache file missing/corrupt */
#define DVC_ENOMEM (-4)   /* device or host allocation failed */
#define DVC_ESTATE (-5)   /* call out of order (e.g. frame state not set up) */

typedef struct dvc_ctx dvc_ctx;

typedef struct {
  int device;            /* HIP device ordinal (reference: -GPUid, configfile.cpp:143-145) */
  int max_iter;          /* BP iteration cap; the reference's MAX_ITERATIONS is 50
                            (LdpcaDecoder.h:4); <= 0 selects 50 */
  const char *data_dir;  /* directory with reg1584.ladder / reg6336.ladder; NULL = the
                            library's own data directory */
  int bp_threads;        /* threads per BP workgroup for N=6336 (0 = default 1024; see dvc_autotune) */
  int parallel_mode;     /* BP semantics: 0 = beliefPropagation (ParallelMode 0/2, LdpcaDecoder.cpp:462-580);
                            1 = beliefPropagation_OpenMP (ParallelMode 1, :164-306): decisions LLR < 0 and a
                            failure after 4 consecutive iterations with unchanged decisions, compared first
                            against the decisions the previous call left (see dvc_ldpca_decode) */
  int rate_step;         /* rate search: 0 = every code from the start code up (LdpcaDecoder.cpp:674);
                            1 = the DYNAMICSTEP build's growing steps (:19-21, 669-672, 760-768): codes
                            start + 0, 1, 3, 6, 10, 14, ... then the full rate; needs parallel_mode 0 */
  int qcif_ladder;       /* the N = 1584 ladder: 0 = the regular RegDecoder1584.lad (65 codes, column
                            degree 3); 1 = the irregular irReg1584.lad (66 codes, column degrees 2..21),
                            the reference's alternative at main.cpp:84-85.  N = 6336 is always regular. */
} dvc_params;

/* InitCUDA + LDPCA_CUDA_Init_Variable (LDPCA_cuda.cu:36-89, 546-763): selects the
 * device and makes both ladders (N=1584 and N=6336) and their full-rate
 * inverses resident. */
int dvc_ctx_create(const dvc_params *params, dvc_ctx **out);
/* LDPCA_CUDA_Free_Variable and the *_Free_Buffer calls. */
int dvc_ctx_destroy(dvc_ctx *ctx);
const char *dvc_last_error(void);
/* Host-only: checks the ladder packs in data_dir (NULL = the library's own
 * data directory) and writes the cached full-rate inverses (<kind><N>.invh)
 * if they are missing.  Needs no GPU. */
int dvc_prepare_data(const char *data_dir);

/* Kernel timing (no counterpart in the reference, which times whole
 * components with gettimeofday, global.cpp:60-76): with profiling enabled the
 * library records HIP events around each kernel launch on the stream it runs
 * on.  dvc_profile(ctx, 1) resets and enables, dvc_profile(ctx, 0) disables;
 * dvc_profile_get sums the elapsed time of every launch of kernel `name`
 * ("bp_6336", "bp_1584", "ce_llr", "invh", ...) since the reset. */
int dvc_profile(dvc_ctx *ctx, int enable);
int dvc_profile_get(dvc_ctx *ctx, const char *name, double *total_ms, long *launches);

/* Ladder facts for code length N: number of codes and m = syndLength[c]
 * (ReadLdpcaLadderFile_LdpcaDecoder.cpp:182). synd_length may be NULL. */
int dvc_ladder_info(dvc_ctx *ctx, int N, int *num_codes, int *num_edges, int *synd_length);
/* Copies the full-rate inverse (N rows of 4*ceil(N/32) bytes, MSB-first, the
 * Inverse_Matrix_H_Reg<N>.bin layout) and decoder rows of code c (E ints). */
int dvc_ladder_export(dvc_ctx *ctx, int N, int code, int *ir_out, uint8_t *inv_out);

typedef struct {
  int start_level;   /* nCodeStart: first code with rate >= H(X|Y) (LdpcaDecoder.cpp:657-659) */
  int level;         /* accepted code index */
  int rate;          /* m + 8 bits (syndrome + CRC, LdpcaDecoder.cpp:681) */
  int requests;      /* codes tried (level - start_level + 1 with rate_step 0) */
  int iterations;    /* BP iterations summed over the codes tried (sequential search) */
  int used_inverse;  /* 1 when the full-rate GF(2) inverse produced the bitplane */
  float entropy;     /* conditional entropy H(X|Y) used for start_level */
  int iterations_run; /* BP iterations the GPU ran for this bitplane: every code launched,
                         speculative codes above `level` and cancelled ones included
                         (equals `iterations` on the sequential path) */
} dvc_bp_stats;

/* One bitplane, sequential rate search: LdpcaDecoder (LdpcaDecoder.cpp:646-771)
 * with beliefPropagation_CUDA (LDPCA_cuda.cu:972-1031) per code.
 *   LR          N likelihood ratios p0/p1 (host memory), as condBitProb writes them
 *   acc         N/8 accumulated-syndrome bytes, MSB first (WZ bitstream layout)
 *   crc         the encoder's CRC-8 of the bitplane
 *   start_level -1: computed on the device from ComputeCE (LdpcaDecoder.cpp:127-143);
 *               >= 0: used as given
 *   decoded     N output bytes (0/1).  With parallel_mode 1 it is also read on
 *               entry: the decisions carried into the first code, as the
 *               reference's decoded buffer carries them from call to call
 *               (bytes other than 0/1 never match a decision)
 * BP follows the context's parallel_mode semantics (0 by default). */
int dvc_ldpca_decode(dvc_ctx *ctx, int N, const float *LR, const uint8_t *acc, uint8_t crc,
                     int start_level, uint8_t *decoded, dvc_bp_stats *stats);

typedef struct {
  int N;                  /* 1584 or 6336 */
  const float *LR;        /* N floats */
  const uint8_t *acc;     /* N/8 bytes */
  uint8_t crc;
  int start_level;        /* -1: from ComputeCE on the device */
  uint8_t *decoded;       /* N bytes out */
  dvc_bp_stats stats;     /* out (iterations = those of the accepted code and the
                             codes below it, as the sequential search would run) */
} dvc_ldpca_item;

/* Many bitplanes of mixed code length in one persistent launch; every code
 * from start_level upward is decoded speculatively and the smallest code
 * whose BP converges and passes the CRC is kept, which is the code the
 * sequential search selects.  Host pointers. */
int dvc_ldpca_decode_batch(dvc_ctx *ctx, int n, dvc_ldpca_item *items);

/* Allocation-free form of dvc_ldpca_decode_batch (no counterpart in the
 * reference; for callers that cannot allocate: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same items and
 * results as dvc_ldpca_decode_batch, except that LR, acc and decoded are
 * device (HBM) pointers.  The call allocates nothing: its working memory is a
 * fixed device array of DVC_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; a batch that needs more (about 150 bitplanes of N = 6336 or
 * 600 of N = 1584 per code length) is rejected with DVC_EINVAL.  Calls that
 * overlap on one GPU share that array and race; the call synchronises the
 * context's stream before it returns.  parallel_mode 0 contexts only. */
#define DVC_NOALLOC_SCRATCH_BYTES (64u << 20)
int dvc_ldpca_decode_batch_noalloc(dvc_ctx *ctx, int n, dvc_ldpca_item *items);

/* BP launch configuration (no counterpart in the reference; the project's
 * autotuner).  dvc_autotune decodes a synthetic calibration batch of
 * `planes` bitplanes per code length (0: 128) with each variant, checks that
 * all variants decode the same bits, keeps the fastest in the context and
 * returns it.  The choice changes time only, never results. */
typedef struct {
  int cif_threads;      /* threads per workgroup for N = 6336: 512, 704 or 1024 */
  int qcif_threads;     /* threads per workgroup for N = 1584: 256 or 512 */
  int qcif_wg_per_cu;   /* QCIF codewords (workgroups) per CU in the persistent grid: 1..4 */
  int poll;             /* BP iterations between a speculative item's cancellation polls: 1..8 */
  double cif_ms, qcif_ms;  /* dvc_autotune: calibration time of the chosen variant */
} dvc_tuning;
int dvc_autotune(dvc_ctx *ctx, int planes, dvc_tuning *out);
int dvc_get_tuning(dvc_ctx *ctx, dvc_tuning *out);
int dvc_set_tuning(dvc_ctx *ctx, const dvc_tuning *tuning);

/* Verification entry (no counterpart in the reference): evaluates one of the
 * device's scalar routines on n inputs.  func 0..3: float logf, expf, tanhf,
 * the BP kernel phi(x) = -logf(tanhf((x+1e-9f)/2)) (float in, float out);
 * 4: double sqrt of int32 inputs; 5: double exp as reconstruction evaluates it, dvc_exp
 * (double in, double out);
 * 6: the branch-free phi for x >= 0 (float in, float out); 7: its shortened
 * form, which the BP kernel evaluates (float in, float out). */
int dvc_selftest_unary(dvc_ctx *ctx, int func, const void *in, void *out, int n);
/* Verification entry (no counterpart in the reference): evaluates the BP
 * kernel's phi and the direct restatement of the reference's
 * -logf(tanhf((x+1e-9f)/2)) on every float bit pattern in [lo, hi] (NaNs
 * skipped).  counts[0] = patterns compared, counts[1] = mismatches,
 * counts[2] = lowest mismatching pattern (0xffffffff when none). */
int dvc_selftest_phi_sweep(dvc_ctx *ctx, uint32_t lo, uint32_t hi, unsigned long long *counts);
/* Calibration entry (no counterpart in the reference): times one launch of a
 * VALU issue-rate kernel -- 8 independent chains per lane of one operation
 * kind (0 fp32 fma, 1 fp32 add, 2 fp32 mul, 3 fp64 fma, 4 int32 add, 5
 * v_sad_u8, 6 v_mqsad_u32_u8, 7 v_qsad_pk_u16_u8, 8 v_rcp_f32), built
 * with the library's own flags -- over waves_per_cu (4..32, a multiple of 4)
 * waves per CU and `iters` trips of 128 operations.  *ms = kernel time (HIP
 * events), *wave_instr = VALU wave64 instructions the launch issues. */
int dvc_selftest_valu_rate(dvc_ctx *ctx, int kind, int waves_per_cu, int iters, double *ms, double *wave_instr);

/* Device buffers for callers without a HIP runtime of their own (no
 * counterpart in the reference): allocation on the context's device and
 * blocking copies on its stream.  kind: 0 host->device, 1 device->host,
 * 2 device->device. */
int dvc_buffer_alloc(dvc_ctx *ctx, size_t bytes, void **dptr);
int dvc_buffer_free(dvc_ctx *ctx, void *dptr);
int dvc_buffer_copy(dvc_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);

/* ---------------------------------------------------------------- WZ frames */
typedef struct {
  int nbitplanes;
  int bitrate;                /* decodedFrames[].bitRate: mode flags + AC ranges + sum of rates */
  int levels[128];            /* accepted code per bitplane (decode order) */
  int start_levels[128];
  int iterations[128];        /* BP iterations of the sequential search per bitplane */
  int used_inverse[128];
  float entropy[128];
  double psnr;                /* against orig when given (frame.cpp:115-124) */
} dvc_frame_stats;

typedef struct {
  int width, height;          /* 176x144 (N = 1584) or 352x288 (N = 6336) */
  int qindex;                 /* 1..8 */
  int past_idx, cur_idx, fut_idx;  /* frame numbers: cur is decoded from past and fut */
  const uint8_t *past, *fut;  /* decoded key frames (luma, W*H) */
  const uint8_t *orig;        /* optional original frame for PSNR */
  const int *ac_range;        /* [16] ACRange of the WZ bitstream (wzbitstream.cpp:19-30) */
  const uint8_t *crc;         /* [nbitplanes] CRC-8 per bitplane */
  const uint8_t *acc;         /* [nbitplanes * N/8] accumulated syndromes */
  uint8_t *recons;            /* out: W*H decoded frame */
  dvc_frame_stats *stats;     /* out, may be NULL */
  /* Block modes (EncodedFrame::BlockModeFlag, wzbitstream.cpp:34-100): one
   * byte per 8x8 block in raster order, 0 = WZ block, 2 = skip block (copied
   * from the past key frame, copySkipBlkToRecons); NULL = every block is a WZ
   * block.  Intra blocks (1) are rejected.  mode_bits = skipBlockBits +
   * intraBlockBits of the frame (the bitrate's starting value,
   * DVC_Decoder.cpp:131); 0 means 2 (both modes off). */
  const uint8_t *block_mode;
  int mode_bits;
  /* Intra blocks (block_mode 1; intraRecon, intra.cpp:240-258): their
   * dequantised coefficients as dvc_wz_stream_frame gives them, 16 per 4x4
   * block (row-major, W/4*H/4 blocks in raster order, zero outside intra
   * blocks); required when block_mode has intra blocks, else NULL. */
  const int32_t *intra_coeffs;
  /* Deblocking (DeblockingMode, DVC_Decoder.cpp:429-435): on/off, the
   * filter's QP (IntraQP), and the past reference's isWZFrame and
   * BlockModeFlag (NULL: all WZ), which set the strength at skip blocks. */
  int deblocking;
  int intra_qp;
  int past_is_wz;
  const uint8_t *past_block_mode;
} dvc_wz_frame;

/* decodeGOP's per-WZ-frame body (DVC_Decoder.cpp:26-474, with skip and
 * intra blocks per block_mode and deblocking per `deblocking`) for n frames.  batched = 0: frames one after another with the
 * sequential rate search per bitplane (the plain drop-in of ParallelMode 2);
 * batched = 1: all frames advance together, one bitplane per step, with the
 * speculative persistent LDPCA launch.  Outputs are identical in both modes
 * and equal the CPU path's (ParallelMode 0). */
int dvc_decode_wz_frames(dvc_ctx *ctx, int n, const dvc_wz_frame *frames, int batched);

/* ---------------------------------------------------------------- WZ bitstream
 * readWZBitstream (wzbitstream.cpp:128-210) and readGOP (wzbitstream.cpp:13-126)
 * over a bitstream held in memory.  The skip and intra block-mode maps are
 * decoded from their JBIG images (replacing simple_JBIG_dec, sJBIG.cpp:49-103,
 * and ByteToBitmap, sJBIG.cpp:24-47, over the vendored JBIG-KIT) by the
 * library's own T.82 decoder. */
typedef struct dvc_wz_stream dvc_wz_stream;
typedef struct {
  int width, height;          /* resolution byte 0: 176x144, 1: 352x288 */
  int qindex, intra_qp, gop, frames, deblocking;
  int wz_frames;              /* WZ frames in the stream */
  int nbitplanes;             /* QbitsNumPerBlk of qindex */
  size_t bytes_read;          /* bytes consumed by the reader */
} dvc_wz_stream_info;
typedef struct {
  int past_idx, cur_idx, fut_idx;  /* begin, (begin + end) / 2, end of the frame's readGOP call */
  int ac_range[16];
  int skip_mode, intra_mode;
  int mode_bits;              /* skipBlockBits + intraBlockBits as readGOP leaves them */
  const uint8_t *block_mode;  /* W/8 * H/8 BlockModeFlag bytes (0 WZ, 1 intra, 2 skip), owned by the stream */
  const uint8_t *crc;         /* [nbitplanes] */
  const uint8_t *acc;         /* [nbitplanes * N/8] accumulated syndromes */
  const int32_t *intra_coeffs;  /* dvc_wz_frame::intra_coeffs once the intra bitstream is attached, else NULL */
  int past_is_wz;             /* the past reference is itself a WZ frame of the stream */
  const uint8_t *past_block_mode;  /* its block modes, or NULL */
} dvc_wz_stream_frame;
/* Parses the whole stream (header, then every GOP's WZ frames in readGOP's
 * hierarchical order, which is decodeGOP's decoding order). */
int dvc_wz_stream_open(const uint8_t *data, size_t len, dvc_wz_stream **out);
int dvc_wz_stream_info_get(const dvc_wz_stream *s, dvc_wz_stream_info *info);
/* The k-th WZ frame of the stream (0 <= k < wz_frames); its pointers stay
 * valid until dvc_wz_stream_close. */
int dvc_wz_stream_frame_get(const dvc_wz_stream *s, int k, dvc_wz_stream_frame *frame);
int dvc_wz_stream_close(dvc_wz_stream *s);
/* The intra bitstream of the stream's frames with intra blocks
 * (readIntraBitstream, intraBitstream.cpp:10-35): parses every such frame's
 * CAVLC blocks in decoding order (read_one_block, intra.cpp:185-215, and
 * readCoeff4x4_CAVLC, vlc.cpp:613-761), which fills their intra_coeffs and
 * adds the bits each frame used to its mode_bits (DVC_Decoder.cpp:43-47). */
int dvc_wz_stream_attach_intra(dvc_wz_stream *s, const uint8_t *data, size_t len);
/* One frame's intra blocks from the intra bitstream, as intraRecon's parsing
 * half does it (read_one_block, intra.cpp:185-215, over readCoeff4x4_CAVLC,
 * vlc.cpp:613-761): the CAVLC blocks of every 4x4 block whose 8x8 block is
 * intra in `block_mode`, in raster order, from bit *bit_offset (advanced
 * past them, as IntraBS.frame_bitoffset is), dequantised with intra_qp into
 * coeffs (16 ints per 4x4 block, W/4*H/4 blocks; zero elsewhere). */
int dvc_intra_parse(const uint8_t *data, size_t len, size_t *bit_offset, const uint8_t *block_mode, int width,
                    int height, int intra_qp, int32_t *coeffs);

/* One block-mode map as readGOP decodes it (simple_JBIG_dec + ByteToBitmap,
 * sJBIG.cpp:24-103): `data` holds the map's JBIG data (no header: the header
 * is the one wzbitstream.cpp:43-52 rebuilds -- width x height, one stripe,
 * TPDON|DPON, MX 8); writes width*height bytes 0/1 to `map` and the data
 * bytes the decoder consumed (through the stripe's end marker) to *consumed. */
int dvc_jbig_decode_map(const uint8_t *data, size_t len, int width, int height, uint8_t *map, size_t *consumed);
/* A whole JBIG bi-level image entity (20-byte header + data) of one plane
 * and the lowest resolution layer (no counterpart in the reference; the
 * general form of the above): the image (one byte 0/1 per pixel, row-major)
 * into `image` when it is not NULL and holds capacity >= width*height bytes. */
int dvc_jbig_decode(const uint8_t *bie, size_t len, uint8_t *image, size_t capacity, uint32_t *width,
                    uint32_t *height, size_t *consumed);

/* ---------------------------------------------------------------- side information
 * createSideInfo (sideInfoCreation.cpp:438-551: key-frame low-pass and
 * half-pel interpolation, forward ME, 16x16 and 8x8 bidirectional half-pel
 * refinement, weighted vector-median smoothing, motion compensation), the
 * residue's forward transform (transform.cpp:153-190) and noiseParameter
 * (noiseModel.cpp:9-75) -- steps 2-4 of decodeGOP, DVC_Decoder.cpp:51-80 --
 * for n frames in one batch.  Replaces forwardME_CUDA (forwardME_kernel.cu:
 * 187-234) and the CPU steps around it.  Outputs equal the CPU path's. */
typedef struct {
  int width, height;          /* 176x144 or 352x288 */
  int qindex;                 /* 1..8: the bands that get alpha (QtableBits > 0) */
  int past_idx, cur_idx, fut_idx;
  const uint8_t *past, *fut;  /* key frames, luma W*H */
  uint8_t *si;                /* out: W*H side information */
  int *residue;               /* out (may be NULL): W*H motion-compensated residue (F - P) >> 1 */
  float *alpha;               /* out (may be NULL): W*H Laplacian parameter at the coefficient
                                 positions of the bands, 0 elsewhere */
} dvc_si_request;
int dvc_side_info_batch(dvc_ctx *ctx, int n, const dvc_si_request *reqs);

/* ---------------------------------------------------------------- per-band entry points
 * The accelerator calls of a DvcDec that keeps its own decodeGOP loop
 * (ParallelMode 2/3 at DVC_Decoder.cpp:91-95, 184-185, 218, 328-329, 380 and
 * sideInfoCreation.cpp:503), on the reference's own data layouts.  Where the
 * reference keeps implicit device state between calls (the band's noise
 * distribution between noiseDistribution_CUDA and condBitProb_CUDA, the key
 * frames uploaded by ME_CUDA_Init_Variable) the context keeps it.  Host
 * pointers; every call returns when its outputs are written.  The SIR lists
 * the WZ 4x4 blocks in raster order, as initSIR (sideInfoCreation.cpp:554-596)
 * does: skip and intra blocks (whole 8x8 blocks) are left out, and the SMF
 * arrays and the distribution are per SIR entry, as the reference's are. */
typedef struct {                 /* SMF, DvcDec/inc/global.h:146-152 */
  float prob[81];
  int searchTop, searchLeft, searchRight, searchBottom;
  int searchWidth, blkIdx, probLen;
} dvc_smf;
typedef struct {                 /* SearchInfo, global.h:158-167 */
  int searchTop, searchBottom, searchLeft, searchRight;
  int top, left, searchWidth, searchRange, blkIdx;
} dvc_search_info;
typedef struct { double x, y; } dvc_point;                  /* Point, global.h:118-120 */
typedef struct { dvc_point past, future, vector; } dvc_mv;  /* MV, global.h:122-126 */

/* CNM_CUDA_Init_Variable (noiseModel_kernel.cu:61) + ME_CUDA_Init_Variable
 * (updateSI_kernel.cu:90): the WZ frame the following per-band calls work on
 * -- its size, Qindex, the decoded key frames past (frame past_idx) and fut
 * (fut_idx), uploaded and filtered once (3x3 low-pass and half-pel FIR), and
 * its SIR (sir may be NULL: every block). */
int dvc_band_frame_init(dvc_ctx *ctx, int width, int height, int qindex, const uint8_t *past, const uint8_t *fut,
                        int past_idx, int cur_idx, int fut_idx, const dvc_search_info *sir, int search_candidate);
/* init_noiseDst_CUDA + noiseDistribution_CUDA (noiseModel_kernel.cu:84, 140;
 * noiseDistribution, noiseModel.cpp:77-132): the band's noise distribution
 * from alpha (W*H), the SMF (one per SIR entry) and the SI's over-complete
 * transform (Trans oct[(H-3)*(W-3)], 16 ints each), kept in the context for
 * dvc_cond_bit_prob.  dist (may be NULL) also receives it in noiseDistribution's
 * layout: per SIR entry `range` (DC) or 2*range+1 (AC) floats. */
int dvc_noise_distribution(dvc_ctx *ctx, int range, int index_in_blk, const float *alpha, const dvc_smf *smf,
                           const int *oct, float *dist);
/* condBitProb_CUDA (noiseModel_kernel.cu:346; condBitProb, noiseModel.cpp:191-310):
 * the N likelihood ratios of bit `bit` of coefficient index_in_blk from the
 * decoded indices so far (recons_trans, W*H) and the distribution of the last
 * dvc_noise_distribution call for the same band (else DVC_ESTATE).  Blocks
 * outside the SIR get 0.9999f / 0.0001f, the value decodeGOP gives them
 * before the call when skip or intra blocks exist (DVC_Decoder.cpp:203-207). */
int dvc_cond_bit_prob(dvc_ctx *ctx, float *LR, int index_in_blk, int bit, int step_size, const int *recons_trans,
                      int range);
/* updateSI_CUDA (updateSI_kernel.cu:298; updateSI + subPelRefine,
 * sideInfoCreation.cpp:598-736): refreshes side_info (W*H, in/out) from the
 * partially reconstructed frame and the key frames of dvc_band_frame_init. */
int dvc_update_si(dvc_ctx *ctx, uint8_t *side_info, const uint8_t *recons);
/* updateSMF_CUDA (motionLearning_kernel.cu:91; updateSMF, motionLearning.cpp:
 * 70-127): refreshes prob[0..probLen) of every SMF (in/out). */
int dvc_update_smf(dvc_ctx *ctx, dvc_smf *smf, const uint8_t *recons, const uint8_t *side_info);
/* forwardME_CUDA (forwardME_kernel.cu:187; forwardME, sideInfoCreation.cpp:
 * 61-145): the macroblock MVs (W/16 * H/16, raster order) of the WZ frame
 * cur_idx from the low-pass key frames.  Independent of dvc_band_frame_init. */
int dvc_forward_me(dvc_ctx *ctx, int width, int height, const uint8_t *fut_lowpass, const uint8_t *past_lowpass,
                   int past_idx, int cur_idx, int fut_idx, dvc_mv *mb_mv);

/* ---------------------------------------------------------------- serving engine
 * Continuous batching (no counterpart in the reference, which decodes frames
 * strictly one after another, main.cpp:132-142): requests are queued by
 * dvc_engine_submit and admitted between steps while their device footprint
 * fits the HBM budget; every dvc_engine_step advances each in-flight frame by
 * at least one bitplane (setting up newcomers first) and retires finished
 * frames, whose recons/stats are then written.  Key frames are kept resident under a
 * caller-chosen key (e.g. stream id << 32 | frame number; 0 = the host
 * pointer) and shared by the frames that name it, until released or evicted
 * (least recently used, only when unreferenced).  The request's host buffers
 * (past, fut, orig, recons, stats) must stay valid until its ticket is done;
 * ac_range/crc/acc are copied at submit (acc at admission when it is a device
 * buffer).  Each frame's result is the same as when decoded alone. */
typedef struct dvc_engine dvc_engine;
typedef struct {
  uint64_t hbm_budget_bytes;  /* 0: 80% of the free device memory */
  int max_inflight;           /* frames decoded concurrently (0: 64) */
  int sequential_bp;          /* 1: plain per-code rate search instead of the speculative launch */
  int lanes;                  /* frame groups stepped concurrently, each on its own stream and host
                                 thread (0: one per 16 frames in flight, at most 4) */
  int eager_steps;            /* 1: launch every step's kernels directly; 0: steps in which no frame
                                 starts or ends a band replay a captured HIP graph of their launches */
  int device_buffers;         /* 1: past, fut, acc and recons of submitted frames are device (HBM)
                                 pointers; ac_range, crc, orig and stats stay host pointers */
  int lane_steps;             /* bitplane steps a lane may take in one dvc_engine_step without
                                 waiting for the other lanes: it stops early when one of its frames
                                 completes, and takes one step while frames wait for admission
                                 (0 or 1: every lane takes one step per call) */
} dvc_engine_params;
int dvc_engine_create(dvc_ctx *ctx, const dvc_engine_params *params, dvc_engine **out);
int dvc_engine_destroy(dvc_engine *eng);
int dvc_engine_submit(dvc_engine *eng, const dvc_wz_frame *frame, uint64_t past_key, uint64_t fut_key,
                      uint64_t *ticket);
int dvc_engine_step(dvc_engine *eng, int *completed);
int dvc_engine_status(dvc_engine *eng, int *queued, int *inflight, uint64_t *hbm_used);
int dvc_engine_done(dvc_engine *eng, uint64_t ticket, int *done);
int dvc_engine_release_key(dvc_engine *eng, uint64_t key);

#ifdef __cplusplus
}
#endif
#endif
