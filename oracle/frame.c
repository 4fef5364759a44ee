/* Whole-frame parts of the oracle: the WZ encoder restatement that produces
 * the decoder's inputs, and decodeGOP's per-WZ-frame body.  TEST
 * INFRASTRUCTURE ONLY (see dvc_oracle.h). */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>
#include "dvc_oracle.h"

static double now_ms(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

int orc_bitplanes_per_frame(int q) {
  int n = 0;
  for (int i = 0; i < 16; i++) n += orc_qtable_bits[q][i];
  return n;
}

/* DvcEnc: forwardTransform (transform.cpp:229-276), quantize
 * (quantization.cpp:95-140, quantize44 :80-91), extractBitplane
 * (bitplane.cpp:4-38) and LdpcaEncoder per bitplane (DVC_Encoder.cpp:87-96),
 * with skip and intra blocks disabled. */
int orc_encode_wz_frame(const orc_geom *g, int q, const orc_enc_ladder *enc, const uint8_t *frame,
                        int *acRange, uint8_t *crc, uint8_t *acc) {
  int W = g->W, N = g->b44n;
  const int *qt = orc_qtable[q], *qb = orc_qtable_bits[q];
  int *trans = (int *)malloc(sizeof(int) * g->frameSize);
  for (int i = 0; i < 16; i++) acRange[i] = 0;
  for (int i = 0; i < g->b88h; i++)
    for (int j = 0; j < g->b88w; j++)
      for (int ii = 0; ii < 8; ii += 4)
        for (int jj = 0; jj < 8; jj += 4) {
          int base = ((i << 3) + ii) * W + (j << 3) + jj, blk[16];
          for (int m = 0; m < 4; m++)
            for (int n = 0; n < 4; n++) blk[4 * m + n] = (int)frame[base + m * W + n];
          orc_forward44(blk);
          for (int m = 0; m < 4; m++)
            for (int n = 0; n < 4; n++) {
              int idx = 4 * m + n;
              trans[base + m * W + n] = blk[idx];
              if (idx != 0 && abs(blk[idx]) > acRange[idx]) acRange[idx] = abs(blk[idx]);
            }
        }
  int step[16], shift[16] = {0};
  step[0] = 1024 / qt[0];
  for (int i = 1; i < 16; i++) {
    step[i] = qt[i] > 0 ? (int)ceil((double)((acRange[i] + 1) << 1) / (double)qt[i]) : 0;
    if (qt[i] > 0) shift[i] = (qt[i] - 2) / 2;
  }
  for (int y = 0; y < g->H; y++)
    for (int x = 0; x < W; x++) {
      int idx = (y & 3) * 4 + (x & 3), *v = &trans[y * W + x];
      if (idx == 0) *v /= step[0];
      else *v = step[idx] > 0 ? *v / step[idx] + shift[idx] : 0;
    }
  int nbp = orc_bitplanes_per_frame(q);
  uint8_t *bp = (uint8_t *)malloc((size_t)nbp * N);
  int bit = 0;
  for (int i = 0; i < g->b44h; i++)
    for (int j = 0; j < g->b44w; j++) {
      int top = i << 2, left = j << 2, k = 0;
      for (int band = 0; band < 16; band++) {
        int index = orc_zigzag[band];
        if (qb[index] <= 0) break;
        int c = trans[(top + (index >> 2)) * W + left + (index & 3)];
        for (int n = qb[index] - 1; n >= 0; n--) bp[(size_t)(k++) * N + bit] = (uint8_t)((c >> n) & 1);
      }
      bit++;
    }
  for (int k = 0; k < nbp; k++) orc_ldpca_encode(enc, bp + (size_t)k * N, acc + (size_t)k * (N / 8), &crc[k]);
  free(bp);
  free(trans);
  return nbp;
}

/* decodeGOP (DVC_Decoder.cpp:26-474), one WZ frame, skip/intra/deblock off. */
int orc_decode_wz_frame(const orc_geom *g, int q, const orc_ladder *L, const uint8_t *past,
                        const uint8_t *fut, int pi, int ci, int fi, const int *acRange,
                        const uint8_t *crc, const uint8_t *acc, const uint8_t *orig,
                        const orc_decode_opts *opt, uint8_t *recons, orc_frame_result *res) {
  int W = g->W, H = g->H, FS = g->frameSize, N = g->b44n;
  const int *qt = orc_qtable[q], *qb = orc_qtable_bits[q];
  int omp = opt->mode == 1;
  memset(res, 0, sizeof(*res));
  double t0 = now_ms(), t;

  uint8_t *plp = (uint8_t *)malloc(FS), *flp = (uint8_t *)malloc(FS);
  uint8_t *phalf = (uint8_t *)malloc(4 * FS), *fhalf = (uint8_t *)malloc(4 * FS);
  uint8_t *si = (uint8_t *)malloc(FS);
  int *residue = (int *)malloc(sizeof(int) * FS), *sideTrans = (int *)malloc(sizeof(int) * FS);
  int *reconsTrans = (int *)calloc(FS, sizeof(int));
  float *alpha = (float *)calloc(FS, sizeof(float));
  orc_mv *mbmv = (orc_mv *)malloc(sizeof(orc_mv) * g->b1616n), *bmv = (orc_mv *)malloc(sizeof(orc_mv) * g->b88n);
  orc_search_info *sir = (orc_search_info *)malloc(sizeof(orc_search_info) * N);
  orc_smf *smf = (orc_smf *)malloc(sizeof(orc_smf) * N);
  int *oct = (int *)malloc(sizeof(int) * 16 * (size_t)(W - 3) * (H - 3));
  float *LR = (float *)malloc(sizeof(float) * N);
  uint8_t *decoded = (uint8_t *)malloc(N), *accBits = (uint8_t *)malloc(N);
  memset(decoded, 0xAA, N); /* content of the fresh malloc under MALLOC_PERTURB_=0x55 */

  t = now_ms();
  orc_lowpass(past, W, H, plp);
  orc_lowpass(fut, W, H, flp);
  orc_fir(past, W, H, phalf);
  orc_fir(fut, W, H, fhalf);
  orc_create_side_info(g, plp, flp, phalf, fhalf, pi, ci, fi, si, residue, mbmv, bmv, omp);
  res->t_si = now_ms() - t;

  orc_forward_transform(g, residue, sideTrans);
  orc_noise_parameter(g, q, sideTrans, alpha, NULL);
  int nsir = orc_init_sir(g, NULL, sir);
  orc_init_smf(g, smf, sir, nsir);
  t = now_ms();
  orc_overcomplete(g, si, oct, omp);
  res->t_oct += now_ms() - t;
  orc_init_si_ml(g, sideTrans, oct);

  int step[16], crange[16];
  res->bitrate = 2; /* skip-mode and intra-mode flags (DVC_Decoder.cpp:131) */
  step[0] = 1024 / qt[0];
  crange[0] = 1024;
  for (int i = 1; i < 16; i++) {
    if (qt[i] > 0) {
      step[i] = (int)ceil((double)((acRange[i] + 1) << 1) / (double)qt[i]);
      res->bitrate += 10;
      crange[i] = ((qt[i] - 2) / 2 + 1) * step[i] - 1;
    } else {
      step[i] = 0;
      crange[i] = 0;
    }
  }
  float *dist = NULL;
  int bpi = 0;
  for (int band = 0; band < 16; band++) {
    int index = orc_zigzag[band], qbits = qb[index];
    if (qbits <= 0) break;
    int range = crange[index];
    t = now_ms();
    size_t per = index == 0 ? (size_t)range : (size_t)(2 * range + 1);
    dist = (float *)realloc(dist, sizeof(float) * per * nsir);
    orc_noise_distribution(g, range, index, alpha, smf, oct, sir, nsir, dist, omp);
    res->t_cnm += now_ms() - t;
    for (int bit = qbits; bit > 0; bit--) {
      t = now_ms();
      orc_cond_bit_prob(g, q, LR, index, bit, step[index], reconsTrans, range, dist, sir, nsir, omp);
      res->t_condbit += now_ms() - t;
      t = now_ms();
      float ent = opt->mode == 0 ? orc_compute_ce(LR, N) : orc_compute_ce_omp(LR, N, opt->nthreads);
      orc_bytes_to_bits(acc + (size_t)bpi * (N / 8), N, accBits);
      orc_ldpca_stats st;
      orc_ldpca_decode_steps(L, LR, accBits, crc[bpi], ent, opt->mode, opt->max_iter, opt->nthreads, opt->rate_step,
                             decoded, &st);
      res->bitrate += st.rate;
      res->levels[bpi] = st.level;
      res->start_levels[bpi] = st.start_level;
      res->iters[bpi] = st.total_iters;
      res->entropy[bpi] = ent;
      orc_put_bitplane(g, decoded, reconsTrans, index, NULL);
      bpi++;
      res->t_ldpca += now_ms() - t;
    }
    orc_reconstruction(g, q, reconsTrans, sideTrans, index, alpha, step, NULL);
    orc_inverse_transform(g, sideTrans, recons, sir, nsir);
    t = now_ms();
    orc_update_si(g, recons, past, fut, phalf, fhalf, pi, ci, fi, sir, nsir, si, NULL, omp);
    res->t_updsi += now_ms() - t;
    t = now_ms();
    orc_overcomplete(g, si, oct, omp);
    res->t_oct += now_ms() - t;
    t = now_ms();
    orc_update_smf(g, smf, recons, si, sir, nsir, omp);
    orc_update_si_ml(g, sideTrans, oct, band + 1, smf, sir, nsir, omp);
    res->t_ml += now_ms() - t;
  }
  orc_inverse_transform(g, sideTrans, recons, sir, nsir);
  res->nbitplanes = bpi;
  res->psnr = orig ? orc_psnr(recons, orig, FS) : 0.0;
  res->t_total = now_ms() - t0;
  free(plp); free(flp); free(phalf); free(fhalf); free(si); free(residue); free(sideTrans);
  free(reconsTrans); free(alpha); free(mbmv); free(bmv); free(sir); free(smf); free(oct);
  free(LR); free(decoded); free(accBits); free(dist);
  return 0;
}
