/* Transform, correlation-noise-model, motion-learning and reconstruction
 * parts of the oracle.  TEST INFRASTRUCTURE ONLY (see dvc_oracle.h).
 * File references are to DvcDec/src/ unless stated. */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "dvc_oracle.h"

#define ORC_MU 0.01f /* defines.h:12 */
#define ORC_ROUND(x) ((x) >= 0.0 ? (int)((x) + 0.5) : -(int)(-(x) + 0.5)) /* defines.h:23 */

/* reconstruction.cpp:9-78 and global.cpp:36-41. */
const int orc_qtable[9][16] = {
    {0},
    {16, 8, 0, 0, 8, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {32, 8, 0, 0, 8, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {32, 8, 4, 0, 8, 4, 0, 0, 4, 0, 0, 0, 0, 0, 0, 0},
    {32, 16, 8, 4, 16, 8, 4, 0, 8, 4, 0, 0, 4, 0, 0, 0},
    {32, 16, 8, 4, 16, 8, 4, 4, 8, 4, 4, 0, 4, 4, 0, 0},
    {64, 16, 8, 8, 16, 8, 8, 4, 8, 8, 4, 4, 8, 4, 4, 0},
    {64, 32, 16, 8, 32, 16, 8, 4, 16, 8, 4, 4, 8, 4, 4, 0},
    {128, 64, 32, 16, 64, 32, 16, 8, 32, 16, 8, 4, 16, 8, 4, 0}};
const int orc_qtable_bits[9][16] = {
    {0},
    {4, 3, 0, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {5, 3, 0, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {5, 3, 2, 0, 3, 2, 0, 0, 2, 0, 0, 0, 0, 0, 0, 0},
    {5, 4, 3, 2, 4, 3, 2, 0, 3, 2, 0, 0, 2, 0, 0, 0},
    {5, 4, 3, 2, 4, 3, 2, 2, 3, 2, 2, 0, 2, 2, 0, 0},
    {6, 4, 3, 3, 4, 3, 3, 2, 3, 3, 2, 2, 3, 2, 2, 0},
    {6, 5, 4, 3, 5, 4, 3, 2, 4, 3, 2, 2, 3, 2, 2, 0},
    {7, 6, 5, 4, 6, 5, 4, 3, 5, 4, 3, 2, 4, 3, 2, 0}};
const int orc_zigzag[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};

static double post_scale[16], pre_scale[16];
static int scales_ready = 0;
/* initTransform (transform.cpp:12-61). */
static void init_scales(void) {
  if (scales_ready) return;
  double a = 0.5, b = sqrt(0.4), a2 = 0.25, b2 = 0.4, ab = a * b;
  double ps[16] = {a2, ab / 2, a2, ab / 2, ab / 2, b2 / 4, ab / 2, b2 / 4,
                   a2, ab / 2, a2, ab / 2, ab / 2, b2 / 4, ab / 2, b2 / 4};
  double pr[16] = {a2, ab, a2, ab, ab, b2, ab, b2, a2, ab, a2, ab, ab, b2, ab, b2};
  for (int i = 0; i < 16; i++) { post_scale[i] = ps[i]; pre_scale[i] = pr[i] * 64.0; }
  scales_ready = 1;
}

/* forward44 (transform.cpp:63-104): rows of the integer core transform are
 * applied down the columns first, then across the rows. */
void orc_forward44(int *blk) {
  init_scales();
  int t[16];
  for (int c = 0; c < 4; c++) {
    int x0 = blk[c], x1 = blk[4 + c], x2 = blk[8 + c], x3 = blk[12 + c];
    t[c] = x0 + x1 + x2 + x3;
    t[4 + c] = (x0 << 1) + x1 - x2 - (x3 << 1);
    t[8 + c] = x0 - x1 - x2 + x3;
    t[12 + c] = x0 - (x1 << 1) + (x2 << 1) - x3;
  }
  for (int r = 0; r < 4; r++) {
    int x0 = t[4 * r], x1 = t[4 * r + 1], x2 = t[4 * r + 2], x3 = t[4 * r + 3];
    blk[4 * r] = x0 + x1 + x2 + x3;
    blk[4 * r + 1] = (x0 << 1) + x1 - x2 - (x3 << 1);
    blk[4 * r + 2] = x0 - x1 - x2 + x3;
    blk[4 * r + 3] = x0 - (x1 << 1) + (x2 << 1) - x3;
  }
  for (int i = 0; i < 16; i++) blk[i] = ORC_ROUND(((double)blk[i]) * post_scale[i]);
}

/* inverse44 (transform.cpp:106-151). */
void orc_inverse44(int *blk) {
  init_scales();
  int t[16];
  for (int i = 0; i < 16; i++) blk[i] = ORC_ROUND(((double)blk[i]) * pre_scale[i]);
  for (int c = 0; c < 4; c++) {
    int x0 = blk[c], x1 = blk[4 + c], x2 = blk[8 + c], x3 = blk[12 + c];
    t[c] = x0 + x1 + x2 + (x3 >> 1);
    t[4 + c] = x0 + (x1 >> 1) - x2 - x3;
    t[8 + c] = x0 - (x1 >> 1) - x2 + x3;
    t[12 + c] = x0 - x1 + x2 - (x3 >> 1);
  }
  for (int r = 0; r < 4; r++) {
    int x0 = t[4 * r], x1 = t[4 * r + 1], x2 = t[4 * r + 2], x3 = t[4 * r + 3];
    blk[4 * r] = x0 + x1 + x2 + (x3 >> 1);
    blk[4 * r + 1] = x0 + (x1 >> 1) - x2 - x3;
    blk[4 * r + 2] = x0 - (x1 >> 1) - x2 + x3;
    blk[4 * r + 3] = x0 - x1 + x2 - (x3 >> 1);
  }
  for (int i = 0; i < 16; i++) blk[i] >>= 6;
}

/* forwardTransform (transform.cpp:153-190). */
void orc_forward_transform(const orc_geom *g, const int *pix, int *trans) {
  int W = g->W;
  for (int i = 0; i < g->b44h; i++)
    for (int j = 0; j < g->b44w; j++) {
      int base = (i << 2) * W + (j << 2), blk[16];
      for (int m = 0; m < 4; m++)
        for (int n = 0; n < 4; n++) blk[4 * m + n] = pix[base + m * W + n];
      orc_forward44(blk);
      for (int m = 0; m < 4; m++)
        for (int n = 0; n < 4; n++) trans[base + m * W + n] = blk[4 * m + n];
    }
}

/* inverseTransform (transform.cpp:234-273). */
void orc_inverse_transform(const orc_geom *g, const int *trans, uint8_t *recons,
                           const orc_search_info *sir, int nsir) {
  int W = g->W;
  for (int i = 0; i < nsir; i++) {
    int base = sir[i].top * W + sir[i].left, blk[16];
    for (int m = 0; m < 4; m++)
      for (int n = 0; n < 4; n++) blk[4 * m + n] = trans[base + m * W + n];
    orc_inverse44(blk);
    for (int m = 0; m < 4; m++)
      for (int n = 0; n < 4; n++) {
        int c = blk[4 * m + n];
        recons[base + m * W + n] = (uint8_t)(c < 0 ? 0 : (c > 255 ? 255 : c));
      }
  }
}

/* overCompleteTransform (transform.cpp:275-298). */
void orc_overcomplete(const orc_geom *g, const uint8_t *pix, int *oct, int omp) {
  int W = g->W, ow = W - 3, oh = g->H - 3;
#pragma omp parallel for if (omp)
  for (int top = 0; top < oh; top++)
    for (int left = 0; left < ow; left++) {
      int *b = oct + 16 * ((size_t)top * ow + left);
      for (int m = 0; m < 4; m++)
        for (int n = 0; n < 4; n++) b[4 * m + n] = pix[(top + m) * W + left + n];
      orc_forward44(b);
    }
}

/* noiseParameter (noiseModel.cpp:9-75). */
void orc_noise_parameter(const orc_geom *g, int qindex, int *trans, float *alpha, const uint8_t *flag) {
  const int *qb = orc_qtable_bits[qindex];
  int W = g->W;
  float mean[17] = {0}, var[17] = {0};
  for (int i = 0; i < g->b44h; i++) {
    int top = i << 2;
    for (int j = 0; j < g->b44w; j++) {
      int left = j << 2, band = 0, index = orc_zigzag[0];
      while (band < 16 && qb[index] > 0) {
        int ci = (top + (index >> 2)) * W + left + (index & 3);
        float c = fabsf((float)trans[ci]);
        mean[band] += c;
        var[band] += c * c;
        trans[ci] = (int)c;
        band++;
        index = orc_zigzag[band < 16 ? band : 15];
      }
    }
  }
  int band = 0, index = orc_zigzag[0];
  while (band < 16 && qb[index] > 0) {
    mean[band] /= (float)g->b44n;
    var[band] = var[band] / (float)g->b44n - mean[band] * mean[band];
    band++;
    index = orc_zigzag[band < 16 ? band : 15];
  }
  for (int i = 0; i < g->b44h; i++) {
    int top = i << 2, row88 = (i >> 1) * g->b88w;
    for (int j = 0; j < g->b44w; j++) {
      if (flag && flag[row88 + (j >> 1)] != 0) continue;
      int left = j << 2;
      band = 0; index = orc_zigzag[0];
      while (band < 16 && qb[index] > 0) {
        int ci = (top + (index >> 2)) * W + left + (index & 3);
        float dst = (float)trans[ci] - mean[band];
        float dst2 = dst * dst;
        if (dst2 > var[band]) alpha[ci] = sqrtf(2 / dst2);
        else alpha[ci] = sqrtf(2 / var[band]);
        band++;
        index = orc_zigzag[band < 16 ? band : 15];
      }
    }
  }
}

/* initSMF (motionLearning.cpp:6-47). */
void orc_init_smf(const orc_geom *g, orc_smf *smf, const orc_search_info *sir, int nsir) {
  int right = g->W - 4, bot = g->H - 4;
  for (int i = 0; i < nsir; i++) {
    int top = sir[i].top, left = sir[i].left;
    int st = top - 4 < 0 ? 0 : top - 4, sb = top + 4 > bot ? bot : top + 4;
    int sl = left - 4 < 0 ? 0 : left - 4, sr = left + 4 > right ? right : left + 4;
    int p = 0;
    for (int m = st; m <= sb; m++)
      for (int n = sl; n <= sr; n++) {
        if (m == top && n == left) { smf[i].blkIdx = p; smf[i].prob[p++] = 1.0f; }
        else smf[i].prob[p++] = 0.0f;
      }
    smf[i].searchTop = st; smf[i].searchBottom = sb; smf[i].searchLeft = sl; smf[i].searchRight = sr;
    smf[i].searchWidth = sr - sl + 1; smf[i].probLen = p;
  }
}

/* init_si_ML (motionLearning.cpp:49-68). */
void orc_init_si_ml(const orc_geom *g, int *si_ml, const int *oct) {
  int W = g->W, ow = W - 3;
  for (int i = 0; i < g->b44h; i++)
    for (int j = 0; j < g->b44w; j++) {
      int top = i << 2, left = j << 2;
      const int *b = oct + 16 * ((size_t)top * ow + left);
      for (int m = 0; m < 4; m++)
        for (int n = 0; n < 4; n++) si_ml[(top + m) * W + left + n] = b[4 * m + n];
    }
}

/* updateSMF (motionLearning.cpp:70-127). */
void orc_update_smf(const orc_geom *g, orc_smf *smf, const uint8_t *recons, const uint8_t *si,
                    const orc_search_info *sir, int nsir, int omp) {
  int W = g->W;
#pragma omp parallel for if (omp)
  for (int i = 0; i < nsir; i++) {
    int top = sir[i].top, left = sir[i].left;
    float block[16];
    for (int m = 0; m < 4; m++)
      for (int n = 0; n < 4; n++) block[4 * m + n] = (float)recons[(top + m) * W + left + n];
    orc_smf s = smf[i];
    int p = 0;
    float sum = 0.0f;
    for (int m = s.searchTop; m <= s.searchBottom; m++)
      for (int n = s.searchLeft; n <= s.searchRight; n++) {
        float sse = 0.0f;
        int b = 0;
        for (int ii = 0; ii < 4; ii++)
          for (int jj = 0; jj < 4; jj++) {
            float d = block[b++] - (float)si[(m + ii) * W + n + jj];
            sse += d * d;
          }
        s.prob[p] = expf((float)-ORC_MU * sse);
        sum += s.prob[p];
        p++;
      }
    if (sum > 0.0) {
      for (int q = 0; q < p; q++) s.prob[q] = s.prob[q] / sum;
    } else {
      s.prob[s.blkIdx] = 1.0f;
    }
    smf[i] = s;
  }
}

/* update_si_ML (motionLearning.cpp:129-160). */
void orc_update_si_ml(const orc_geom *g, int *sideTrans, const int *oct, int band, const orc_smf *smf,
                      const orc_search_info *sir, int nsir, int omp) {
  int W = g->W, ow = W - 3;
#pragma omp parallel for if (omp)
  for (int i = 0; i < nsir; i++) {
    int top = sir[i].top, left = sir[i].left;
    float block[16] = {0};
    const orc_smf *s = &smf[i];
    int p = 0;
    for (int m = s->searchTop; m <= s->searchBottom; m++)
      for (int n = s->searchLeft; n <= s->searchRight; n++) {
        const int *b = oct + 16 * ((size_t)m * ow + n);
        float prob = s->prob[p];
        for (int k = band; k < 16; k++) {
          int index = orc_zigzag[k];
          block[index] += prob * (float)b[index];
        }
        p++;
      }
    for (int k = band; k < 16; k++) {
      int index = orc_zigzag[k];
      sideTrans[(top + (index >> 2)) * W + left + (index & 3)] = (int)block[index];
    }
  }
}

/* noiseDistribution (noiseModel.cpp:77-132); _OpenMP (:134-188) splits the
 * blocks across threads with the same per-block arithmetic. */
int orc_noise_distribution(const orc_geom *g, int range, int ib, const float *alpha, const orc_smf *smf,
                           const int *oct, const orc_search_info *sir, int nsir, float *dist, int omp) {
  int W = g->W, ow = W - 3;
  if (ib == 0) {
#pragma omp parallel for if (omp)
    for (int i = 0; i < nsir; i++) {
      int top = sir[i].top, left = sir[i].left;
      float a = alpha[top * W + left];
      int side = oct[16 * ((size_t)top * ow + left)];
      float *prob = dist + (size_t)i * range;
      for (int k = 0; k < range; k++) prob[k] = expf(-a * (float)abs(k - side));
    }
    return range;
  }
  int rmax = (range << 1) + 1, br = ib >> 2, bc = ib & 3;
#pragma omp parallel for if (omp)
  for (int i = 0; i < nsir; i++) {
    int top = sir[i].top, left = sir[i].left;
    float a = alpha[(top + br) * W + left + bc];
    const orc_smf *s = &smf[i];
    int side[81], n = 0;
    for (int m = s->searchTop; m <= s->searchBottom; m++)
      for (int q = s->searchLeft; q <= s->searchRight; q++) side[n++] = oct[16 * ((size_t)m * ow + q) + ib];
    float *prob = dist + (size_t)i * rmax;
    for (int k = 0; k < rmax; k++) {
      float sum = 0.0f;
      int pc = k - range;
      for (int m = 0; m < n; m++) sum += s->prob[m] * expf(-a * (float)abs(pc - side[m]));
      prob[k] = sum;
    }
  }
  return rmax;
}

/* condBitProb (noiseModel.cpp:191-310); _OpenMP (:314-436) splits the blocks.
 * Interval indices are clamped to the distribution (they stay inside it for
 * every bitstream the reference encoder produces). */
void orc_cond_bit_prob(const orc_geom *g, int qindex, float *LR, int ib, int bit, int step,
                       const int *rt, int range, const float *dist, const orc_search_info *sir, int nsir,
                       int omp) {
  int W = g->W, iq = (1 << (bit - 1)) - 1;
  const int *qt = orc_qtable[qindex];
  if (ib == 0) {
#pragma omp parallel for if (omp)
    for (int i = 0; i < nsir; i++) {
      int rc = rt[sir[i].top * W + sir[i].left];
      const float *prob = dist + (size_t)i * range;
      int q0l = rc << bit, q1l = ((rc << 1) + 1) << (bit - 1);
      int q0u = q0l + iq, q1u = q1l + iq;
      int v0l = q0l * step, v1l = q1l * step, v0u = (q0u + 1) * step, v1u = (q1u + 1) * step;
      if (v0l < 0) v0l = 0;
      if (v1l < 0) v1l = 0;
      if (v0u > range) v0u = range;
      if (v1u > range) v1u = range;
      float p0 = 0.0f, p1 = 0.0f;
      for (int k = v0l; k < v0u; k++) p0 += prob[k];
      for (int k = v1l; k < v1u; k++) p1 += prob[k];
      p0 = p0 / (p0 + p1);
      p1 = 1 - p0;
      if (p0 <= 0.0001) LR[sir[i].blkIdx] = (float)(0.0001 / 0.9999);
      else if (p0 >= 0.9999) LR[sir[i].blkIdx] = (float)(0.9999 / 0.0001);
      else LR[sir[i].blkIdx] = p0 / p1;
    }
    return;
  }
  int br = ib >> 2, bc = ib & 3, ls = (qt[ib] - 2) / 2, rmax = (range << 1) + 1;
#pragma omp parallel for if (omp)
  for (int i = 0; i < nsir; i++) {
    int rc = rt[(sir[i].top + br) * W + sir[i].left + bc];
    const float *prob = dist + (size_t)i * rmax;
    int q0l = (rc << bit) - ls, q1l = (((rc << 1) + 1) << (bit - 1)) - ls;
    int q0u = q0l + iq, q1u = q1l + iq;
    if (q1u > ls) q1u = ls;
    int v0l = (q0l <= 0) ? (q0l - 1) * step + 1 : q0l * step;
    int v1l = (q1l <= 0) ? (q1l - 1) * step + 1 : q1l * step;
    int v0u = (q0u < 0) ? q0u * step + 1 : (q0u + 1) * step;
    int v1u = (q1u < 0) ? q1u * step + 1 : (q1u + 1) * step;
    v0l += range; v0u += range; v1l += range; v1u += range;
    if (v0l < 0) v0l = 0;
    if (v1l < 0) v1l = 0;
    if (v0u > rmax) v0u = rmax;
    if (v1u > rmax) v1u = rmax;
    float p0 = 0.0f, p1 = 0.0f;
    for (int k = v0l; k < v0u; k++) p0 += prob[k];
    for (int k = v1l; k < v1u; k++) p1 += prob[k];
    p0 = p0 / (p0 + p1);
    p1 = 1 - p0;
    if (p0 <= 0.0001) LR[sir[i].blkIdx] = (float)(0.0001 / 0.9999);
    else if (p0 >= 0.9999) LR[sir[i].blkIdx] = (float)(0.9999 / 0.0001);
    else LR[sir[i].blkIdx] = p0 / p1;
  }
}

/* putBitplaneToCoeff (bitplane.cpp:4-22). */
void orc_put_bitplane(const orc_geom *g, const uint8_t *bp, int *rt, int ib, const uint8_t *flag) {
  int W = g->W, br = ib >> 2, bc = ib & 3, b = 0;
  for (int i = 0; i < g->b44h; i++) {
    int row = ((i << 2) + br) * W, row88 = (i >> 1) * g->b88w;
    for (int j = 0; j < g->b44w; j++) {
      if (!flag || flag[row88 + (j >> 1)] == 0) {
        int p = row + (j << 2) + bc;
        rt[p] = (rt[p] << 1) + (int)bp[b];
      }
      b++;
    }
  }
}

/* reconstruction (reconstruction.cpp:184-289) without the deblocking flags. */
void orc_reconstruction(const orc_geom *g, int qindex, const int *rt, int *st, int ib, const float *alpha,
                        const int *stepSize, const uint8_t *flag) {
  int W = g->W, br = ib >> 2, bc = ib & 3, step = stepSize[ib];
  int ls = ib ? (orc_qtable[qindex][ib] - 2) / 2 : 0;
  for (int i = 0; i < g->b44h; i++) {
    int row = ((i << 2) + br) * W, row88 = (i >> 1) * g->b88w;
    for (int j = 0; j < g->b44w; j++) {
      if (flag && flag[row88 + (j >> 1)] != 0) continue;
      int ci = row + (j << 2) + bc, lo, hi;
      int side = st[ci];
      double a = alpha[ci];
      if (ib == 0) {
        int q = rt[ci];
        lo = q * step; hi = (q + 1) * step;
      } else {
        int q = rt[ci] - ls;
        if (q == 0) { lo = -step + 1; hi = step; }
        else if (q > 0) { lo = q * step; hi = (q + 1) * step; }
        else { lo = (q - 1) * step + 1; hi = q * step + 1; }
      }
      if (a <= 0.0001) {
        st[ci] = (lo + hi) / 2;
      } else if (side < lo) {
        st[ci] = lo + ORC_ROUND(1.0 / a + (double)step / (1.0 - exp(a * (double)step)));
      } else if (side >= hi) {
        st[ci] = hi - ORC_ROUND(1.0 / a + (double)step / (1.0 - exp(a * (double)step)));
      } else {
        double r = side - lo, d = hi - side;
        double er = exp(-a * r), ed = exp(-a * d);
        st[ci] = side + ORC_ROUND(((r + 1.0 / a) * er - (d + 1.0 / a) * ed) / (2.0 - (er + ed)));
      }
    }
  }
}

/* psnr (frame.cpp:115-124). */
double orc_psnr(const uint8_t *a, const uint8_t *b, int n) {
  double mse = 0.0;
  for (int i = 0; i < n; i++) {
    double d = (double)a[i] - (double)b[i];
    mse += d * d;
  }
  mse /= (double)n;
  return mse == 0.0 ? 0.0 : 10.0 * log10(255.0 * 255.0 / mse);
}
