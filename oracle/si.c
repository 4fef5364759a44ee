/* Side-information part of the oracle: key-frame filters, forward and
 * bidirectional motion estimation, weighted vector-median smoothing, motion
 * compensation and the per-band SI refresh.  TEST INFRASTRUCTURE ONLY (see
 * dvc_oracle.h).  File references are to DvcDec/src/. */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "dvc_oracle.h"

#define ORC_K 0.05f      /* defines.h:10 (float literal promoted to double) */
#define ORC_SEARCHRANGE 32

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

void orc_geom_init(orc_geom *g, int W, int H) {
  g->W = W; g->H = H; g->frameSize = W * H;
  g->b44n = g->frameSize / 16; g->b44w = W / 4; g->b44h = H / 4;
  g->b88n = g->b44n / 4; g->b88w = W / 8; g->b88h = H / 8;
  g->b1616n = g->b88n / 4; g->b1616w = W / 16; g->b1616h = H / 16;
}

/* costFunc (sideInfoCreation.cpp:10-12). */
static inline double cost_func(double mad, double x, double y) {
  return mad * (1.0 + ORC_K * sqrt(x * x + y * y));
}
/* computeEndPoint (sideInfoCreation.cpp:14-22). */
static inline int end_point(double st, double vec, double scale, int lower, int upper) {
  int e = (int)((st + vec * scale) + 0.5);
  return e < lower ? lower : (e > upper ? upper : e);
}

/* lowPassFilter (filter.cpp:214-233). */
void orc_lowpass(const uint8_t *in, int W, int H, uint8_t *out) {
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++) {
      double sum = 0;
      for (int m = i - 1; m <= i + 1; m++) {
        if (m < 0 || m >= H) continue;
        for (int n = j - 1; n <= j + 1; n++) {
          if (n < 0 || n >= W) continue;
          sum += (double)in[m * W + n];
        }
      }
      out[i * W + j] = (uint8_t)(sum / 9);
    }
}

/* FIR_filter (filter.cpp:55-117) restricted to the emitted (non-padding)
 * samples; coefficient tables at filter.cpp:30-52. */
void orc_fir(const uint8_t *in, int W, int H, uint8_t *out) {
  static const double tap[6] = {1.0, -5.0, 20.0, 20.0, -5.0, 1.0};
  int W2 = W * 2, H2 = H * 2;
  for (int i = 0; i < H2; i++)
    for (int j = 0; j < W2; j++) {
      int sub = (j & 1) + ((i & 1) << 1);
      if (!sub) { out[i * W2 + j] = in[imin(H - 1, i / 2) * W + imin(W - 1, j / 2)]; continue; }
      double sum = 0.0;
      for (int ii = 0; ii < 6; ii++) {
        int py = imax(0, imin(H - 1, i / 2 - 2 + ii));
        for (int jj = 0; jj < 6; jj++) {
          int px = imax(0, imin(W - 1, j / 2 - 2 + jj));
          double c;
          if (sub == 1) c = (ii == 2) ? tap[jj] / 32.0 : 0.0;        /* a_pos */
          else if (sub == 2) c = (jj == 2) ? tap[ii] / 32.0 : 0.0;   /* b_pos */
          else c = tap[ii] * tap[jj] / 1024.0;                        /* c_pos */
          sum += c * in[py * W + px];
        }
      }
      out[i * W2 + j] = (uint8_t)imax(0, imin(255, (int)(sum + 0.5)));
    }
}

/* init_1616ME (sideInfoCreation.cpp:24-59). */
void orc_init_1616(const orc_geom *g, orc_search_info *s) {
  int bot = g->H - 16, right = g->W - 16, idx = 0;
  for (int i = 0; i < g->b1616h; i++) {
    int top = i << 4, st = imax(0, top - ORC_SEARCHRANGE), sb = imin(bot, top + ORC_SEARCHRANGE);
    for (int j = 0; j < g->b1616w; j++) {
      int left = j << 4, sl = imax(0, left - ORC_SEARCHRANGE), sr = imin(right, left + ORC_SEARCHRANGE);
      orc_search_info v = {st, sb, sl, sr, top, left, sr - sl + 1, (sr - sl + 1) * (sb - st + 1), idx};
      s[idx++] = v;
    }
  }
}

/* forwardME (sideInfoCreation.cpp:61-145); forwardME_OpenMP (:147-232) is the
 * same per-macroblock computation split across threads. */
void orc_forward_me(const orc_geom *g, const orc_search_info *si1616, const uint8_t *fut,
                    const uint8_t *past, int pastIdx, int curIdx, int futIdx, orc_mv *mbmv,
                    orc_point *holes, int omp) {
  int W = g->W, bot = g->H - 16, right = g->W - 16, n = g->b1616n;
  double iFC = (double)(futIdx - curIdx) / (double)(futIdx - pastIdx);
  double iPC = (double)(curIdx - pastIdx) / (double)(futIdx - pastIdx);
  orc_mv *fwd = (orc_mv *)malloc(sizeof(orc_mv) * n);
#pragma omp parallel for if (omp)
  for (int i = 0; i < n; i++) {
    orc_search_info s = si1616[i];
    int fb[256], b = 0;
    for (int m = 0; m < 16; m++)
      for (int q = 0; q < 16; q++) fb[b++] = fut[(s.top + m) * W + s.left + q];
    double minCost = 10000000.0;
    orc_mv best;
    memset(&best, 0, sizeof(best));
    for (int py = s.searchTop; py <= s.searchBottom; py++)
      for (int px = s.searchLeft; px <= s.searchRight; px++) {
        int sad = 0;
        b = 0;
        for (int m = 0; m < 16; m++)
          for (int q = 0; q < 16; q++) sad += abs(fb[b++] - past[(py + m) * W + px + q]);
        double vx = (double)(px - s.left), vy = (double)(py - s.top);
        double cost = cost_func((double)sad, vx, vy);
        if (cost < minCost) {
          minCost = cost;
          best.past.x = (double)px; best.past.y = (double)py;
          best.vector.x = vx; best.vector.y = vy;
        }
      }
    fwd[i] = best;
    holes[i].x = (double)s.left * iPC + (double)best.past.x * iFC;
    holes[i].y = (double)s.top * iPC + (double)best.past.y * iFC;
  }
  double maxDst = W * W + g->H * g->H;
#pragma omp parallel for if (omp)
  for (int i = 0; i < n; i++) {
    int top = si1616[i].top, left = si1616[i].left;
    double minDst = maxDst;
    orc_mv best = fwd[0];
    for (int j = 0; j < n; j++) {
      double dst = (left - holes[j].x) * (left - holes[j].x) + (top - holes[j].y) * (top - holes[j].y);
      if (dst < minDst) { minDst = dst; best = fwd[j]; }
    }
    orc_mv *o = &mbmv[i];
    o->future.x = (double)end_point(left, -best.vector.x, iFC, 0, right);
    o->future.y = (double)end_point(top, -best.vector.y, iFC, 0, bot);
    o->past.x = (double)end_point(left, best.vector.x, iPC, 0, right);
    o->past.y = (double)end_point(top, best.vector.y, iPC, 0, bot);
    o->vector.x = o->past.x - o->future.x;
    o->vector.y = o->past.y - o->future.y;
  }
  free(fwd);
}

/* bidirectME (sideInfoCreation.cpp:235-310): raster order, in place. */
void orc_bidirect_me(const orc_geom *g, const uint8_t *fut, const uint8_t *past, int pastIdx,
                     int curIdx, int futIdx, orc_mv *mv, int blockLen) {
  int bh = g->H / blockLen, bw = g->W / blockLen;
  double iFC = (double)(futIdx - curIdx) / (double)(futIdx - pastIdx);
  double iPC = (double)(curIdx - pastIdx) / (double)(futIdx - pastIdx);
  int W2 = g->W << 1, H2 = g->H << 1, bl2 = blockLen << 1;
  int bot = H2 - bl2, right = W2 - bl2, idx = 0;
  for (int i = 0; i < bh; i++) {
    int y2 = i * bl2;
    for (int j = 0; j < bw; j++) {
      int x2 = j * bl2;
      int sT = (i - 1 >= 0) ? (int)mv[idx - bw].future.y + bl2 : (int)mv[idx].future.y;
      int sB = (i + 1 < bh) ? (int)mv[idx + bw].future.y - bl2 : (int)mv[idx].future.y;
      int sL = (j - 1 >= 0) ? (int)mv[idx - 1].future.x + bl2 : (int)mv[idx].future.x;
      int sR = (j + 1 < bw) ? (int)mv[idx + 1].future.x - bl2 : (int)mv[idx].future.x;
      double minCost = 10000000.0;
      orc_mv best = mv[idx];
      double scale = iPC / iFC;
      for (int fy = sT; fy <= sB; fy++) {
        int py = end_point(y2, (double)(y2 - fy), scale, 0, bot);
        for (int fx = sL; fx <= sR; fx++) {
          int px = end_point(x2, (double)(x2 - fx), scale, 0, right);
          int mad = 0;
          for (int m = 0; m < bl2; m += 2) {
            int pI = (py + m) * W2 + px, fI = (fy + m) * W2 + fx;
            for (int q = 0; q < bl2; q += 2) mad += abs((int)fut[fI + q] - (int)past[pI + q]);
          }
          double vx = (double)(px - fx), vy = (double)(py - fy);
          double cost = cost_func((double)mad, vx, vy);
          if (cost < minCost) {
            minCost = cost;
            best.future.x = (double)fx; best.future.y = (double)fy;
            best.past.x = (double)px; best.past.y = (double)py;
            best.vector.x = vx; best.vector.y = vy;
          }
        }
      }
      mv[idx++] = best;
    }
  }
}

/* motionFilterAndComp (sideInfoCreation.cpp:312-436). */
void orc_motion_filter_comp(const orc_geom *g, const uint8_t *fut, const uint8_t *past, int pastIdx,
                            int curIdx, int futIdx, orc_mv *BlockMV, double *MSE, uint8_t *si,
                            int *residue) {
  int W = g->W, W2 = g->W << 1, H2 = g->H << 1;
  double iFC = (double)(futIdx - curIdx) / (double)(futIdx - pastIdx);
  double iPC = (double)(curIdx - pastIdx) / (double)(futIdx - pastIdx);
  int bot = H2 - 16, right = W2 - 16;
  for (int i = 0; i < g->b88n; i++) {
    int fy = (int)BlockMV[i].future.y, fx = (int)BlockMV[i].future.x;
    int py = (int)BlockMV[i].past.y, px = (int)BlockMV[i].past.x;
    double mse = 0.0;
    for (int m = 0; m < 16; m += 2)
      for (int q = 0; q < 16; q += 2) {
        double d = (double)fut[(fy + m) * W2 + fx + q] - (double)past[(py + m) * W2 + px + q];
        mse += d * d;
      }
    MSE[i] = mse / 64;
  }
  orc_mv *out = (orc_mv *)malloc(sizeof(orc_mv) * g->b88n);
  int idx = 0;
  for (int i = 0; i < g->b88h; i++) {
    int y2 = i << 4;
    for (int j = 0; j < g->b88w; j++) {
      int x2 = j << 4;
      int nb[9], cnt = 0;
      for (int m = i - 1; m <= i + 1; m++) {
        if (m < 0 || m >= g->b88h) continue;
        for (int q = j - 1; q <= j + 1; q++) {
          if (q < 0 || q >= g->b88w) continue;
          nb[cnt++] = m * g->b88w + q;
        }
      }
      double minW = 0.0;
      int best = nb[0];
      for (int k = 1; k < cnt; k++) {
        int o = nb[k];
        double dx = BlockMV[best].vector.x - BlockMV[o].vector.x;
        double dy = BlockMV[best].vector.y - BlockMV[o].vector.y;
        minW += MSE[best] / MSE[o] * sqrt(dx * dx + dy * dy);
      }
      for (int c = 1; c < cnt; c++) {
        int cand = nb[c];
        double w = 0.0;
        for (int k = 0; k < cnt; k++) {
          if (c == k) continue;
          int o = nb[k];
          double dx = BlockMV[cand].vector.x - BlockMV[o].vector.x;
          double dy = BlockMV[cand].vector.y - BlockMV[o].vector.y;
          w += MSE[cand] / MSE[o] * sqrt(dx * dx + dy * dy);
        }
        if (w < minW) { minW = w; best = cand; }
      }
      orc_point v = BlockMV[best].vector;
      orc_mv r;
      r.future.x = (double)end_point(x2, -v.x, iFC, 0, right);
      r.future.y = (double)end_point(y2, -v.y, iFC, 0, bot);
      r.past.x = (double)end_point(x2, v.x, iPC, 0, right);
      r.past.y = (double)end_point(y2, v.y, iPC, 0, bot);
      r.vector.x = r.past.x - r.future.x;
      r.vector.y = r.past.y - r.future.y;
      out[idx++] = r;
    }
  }
  memcpy(BlockMV, out, sizeof(orc_mv) * g->b88n);
  free(out);
  idx = 0;
  for (int i = 0; i < g->b88h; i++) {
    int y = i << 3;
    for (int j = 0; j < g->b88w; j++) {
      int x = j << 3;
      orc_mv v = BlockMV[idx++];
      int fy = (int)v.future.y, fx = (int)v.future.x, py = (int)v.past.y, px = (int)v.past.x;
      for (int m = 0; m < 16; m += 2) {
        int pI = (py + m) * W2 + px, fI = (fy + m) * W2 + fx, cI = (y + (m >> 1)) * W + x;
        for (int q = 0; q < 16; q += 2) {
          si[cI + (q >> 1)] = (uint8_t)(int)((double)fut[fI + q] * iFC + (double)past[pI + q] * iPC);
          residue[cI + (q >> 1)] = ((int)fut[fI + q] - (int)past[pI + q]) >> 1;
        }
      }
    }
  }
}

/* createSideInfo (sideInfoCreation.cpp:438-551). */
void orc_create_side_info(const orc_geom *g, const uint8_t *past_lp, const uint8_t *fut_lp,
                          const uint8_t *past_half, const uint8_t *fut_half, int pastIdx, int curIdx,
                          int futIdx, uint8_t *si, int *residue, orc_mv *mbmv, orc_mv *blockMV,
                          int omp) {
  orc_search_info *s1616 = (orc_search_info *)malloc(sizeof(orc_search_info) * g->b1616n);
  orc_point *holes = (orc_point *)malloc(sizeof(orc_point) * g->b1616n);
  double *mse = (double *)malloc(sizeof(double) * g->b88n);
  orc_init_1616(g, s1616);
  orc_forward_me(g, s1616, fut_lp, past_lp, pastIdx, curIdx, futIdx, mbmv, holes, omp);
  for (int i = 0; i < g->b1616n; i++) {
    orc_mv v = mbmv[i];
    v.future.x *= 2; v.future.y *= 2; v.past.x *= 2; v.past.y *= 2;
    v.vector.x = v.past.x - v.future.x; v.vector.y = v.past.y - v.future.y;
    mbmv[i] = v;
  }
  orc_bidirect_me(g, fut_half, past_half, pastIdx, curIdx, futIdx, mbmv, 16);
  double iFC = (double)(futIdx - curIdx) / (double)(futIdx - pastIdx);
  double iPC = (double)(curIdx - pastIdx) / (double)(futIdx - pastIdx);
  int bot = (g->H - 8) << 1, right = (g->W - 8) << 1, idx = 0;
  for (int i = 0; i < g->b88h; i++) {
    int row = (i >> 1) * g->b1616w, y2 = i << 4;
    for (int j = 0; j < g->b88w; j++) {
      int x2 = j << 4;
      orc_point v = mbmv[row + (j >> 1)].vector;
      orc_mv r;
      r.future.x = (double)end_point(x2, -v.x, iFC, 0, right);
      r.future.y = (double)end_point(y2, -v.y, iFC, 0, bot);
      r.past.x = (double)end_point(x2, v.x, iPC, 0, right);
      r.past.y = (double)end_point(y2, v.y, iPC, 0, bot);
      r.vector.x = r.past.x - r.future.x;
      r.vector.y = r.past.y - r.future.y;
      blockMV[idx++] = r;
    }
  }
  orc_bidirect_me(g, fut_half, past_half, pastIdx, curIdx, futIdx, blockMV, 8);
  orc_motion_filter_comp(g, fut_half, past_half, pastIdx, curIdx, futIdx, blockMV, mse, si, residue);
  free(s1616); free(holes); free(mse);
}

/* initSIR (sideInfoCreation.cpp:554-596).  Returns the candidate count. */
int orc_init_sir(const orc_geom *g, const uint8_t *flag, orc_search_info *sir) {
  int bot = g->H - 4, right = g->W - 4, n = 0, b44 = 0;
  for (int i = 0; i < g->b44h; i++) {
    int top = i << 2, st = imax(0, top - ORC_SEARCHRANGE), sb = imin(bot, top + ORC_SEARCHRANGE);
    int row88 = (i >> 1) * g->b88w;
    for (int j = 0; j < g->b44w; j++) {
      if (flag == NULL || flag[row88 + (j >> 1)] == 0) {
        int left = j << 2, sl = imax(0, left - ORC_SEARCHRANGE), sr = imin(right, left + ORC_SEARCHRANGE);
        orc_search_info v = {st, sb, sl, sr, top, left, sr - sl + 1, (sr - sl + 1) * (sb - st + 1), b44};
        sir[n++] = v;
      }
      b44++;
    }
  }
  return n;
}

/* subPelRefine (sideInfoCreation.cpp:598-640). */
static void sub_pel_refine(const double *block, const uint8_t *ref, int *bx, int *by, int W2, int H2) {
  int bot = H2 - 8, right = W2 - 8, iX = (*bx) << 1, iY = (*by) << 1;
  double minCost = 10000000.0;
  for (int i = -1; i <= 1; i++) {
    int top = iY + i;
    if (top < 0 || top > bot) continue;
    for (int j = -1; j <= 1; j++) {
      int left = iX + j;
      if (left < 0 || left > right) continue;
      double cost = 0.0;
      int b = 0, row = top * W2 + left;
      for (int m = 0; m < 8; m += 2) {
        cost += fabs(block[b++] - (double)ref[row]);
        cost += fabs(block[b++] - (double)ref[row + 2]);
        cost += fabs(block[b++] - (double)ref[row + 4]);
        cost += fabs(block[b++] - (double)ref[row + 6]);
        row += W2 << 1;
      }
      if (cost < minCost) { minCost = cost; *bx = left; *by = top; }
    }
  }
}

/* updateSI (sideInfoCreation.cpp:642-736); updateSI_OpenMP (:738-832) splits
 * the same per-block work across threads. */
void orc_update_si(const orc_geom *g, const uint8_t *recons, const uint8_t *past, const uint8_t *fut,
                   const uint8_t *past_half, const uint8_t *fut_half, int pastIdx, int curIdx, int futIdx,
                   const orc_search_info *sir, int nsir, uint8_t *si, int *mv_out, int omp) {
  int W = g->W, W2 = W << 1, H2 = g->H << 1;
  double iFC = (double)(futIdx - curIdx) / (double)(futIdx - pastIdx);
  double iPC = (double)(curIdx - pastIdx) / (double)(futIdx - pastIdx);
#pragma omp parallel for if (omp) schedule(dynamic, 16)
  for (int i = 0; i < nsir; i++) {
    orc_search_info s = sir[i];
    double block[16];
    int b = 0;
    for (int m = 0; m < 4; m++)
      for (int q = 0; q < 4; q++) block[b++] = (double)recons[(s.top + m) * W + s.left + q];
    int pbx = 0, pby = 0, fbx = 0, fby = 0;
    double pMin = 10000000.0, fMin = 10000000.0;
    for (int y = s.searchTop; y <= s.searchBottom; y++)
      for (int x = s.searchLeft; x <= s.searchRight; x++) {
        double pc = 0.0, fc = 0.0;
        b = 0;
        for (int m = 0; m < 4; m++) {
          int row = (y + m) * W + x;
          for (int q = 0; q < 4; q++) {
            pc += fabs(block[b] - (double)past[row + q]);
            fc += fabs(block[b] - (double)fut[row + q]);
            b++;
          }
        }
        double mx = (double)((x - s.left) << 1), my = (double)((y - s.top) << 1);
        pc = cost_func(pc, mx, my);
        fc = cost_func(fc, mx, my);
        if (pc < pMin) { pbx = x; pby = y; pMin = pc; }
        if (fc < fMin) { fbx = x; fby = y; fMin = fc; }
      }
    sub_pel_refine(block, past_half, &pbx, &pby, W2, H2);
    sub_pel_refine(block, fut_half, &fbx, &fby, W2, H2);
    if (mv_out) { mv_out[4 * i] = pbx; mv_out[4 * i + 1] = pby; mv_out[4 * i + 2] = fbx; mv_out[4 * i + 3] = fby; }
    for (int m = 0; m < 8; m += 2) {
      int row = (s.top + (m >> 1)) * W + s.left, pR = (pby + m) * W2 + pbx, fR = (fby + m) * W2 + fbx;
      for (int q = 0; q < 8; q += 2)
        si[row + (q >> 1)] = (uint8_t)(int)((double)past_half[pR + q] * iPC + (double)fut_half[fR + q] * iFC);
    }
  }
}
