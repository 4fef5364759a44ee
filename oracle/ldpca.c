/* LDPCA part of the oracle: ladder I/O and regeneration, CRC-8, conditional
 * entropy, syndrome formation, sum-product BP (modes 0 and 1), the full-rate
 * GF(2) inverse and the rate search.  TEST INFRASTRUCTURE ONLY (see
 * dvc_oracle.h). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>
#include "dvc_oracle.h"

static int read_ints(FILE *f, int *dst, int n) {
  for (int i = 0; i < n; i++)
    if (fscanf(f, "%d", dst + i) != 1) return -1;
  return 0;
}

/* ReadLadder_decoder (ReadLdpcaLadderFile_LdpcaDecoder.cpp:131-250). */
int orc_read_decoder_ladder(orc_ladder *L, const char *lad_path, const char *inv_path) {
  memset(L, 0, sizeof(*L));
  FILE *f = fopen(lad_path, "r");
  if (!f) return -1;
  if (fscanf(f, "%d %d %d %d", &L->numCodes, &L->N, &L->E, &L->period) != 4) { fclose(f); return -1; }
  int C = L->numCodes, N = L->N, E = L->E, P = L->period;
  L->jc = (int *)malloc(sizeof(int) * (N + 1));
  L->bitsPerPeriod = (int *)malloc(sizeof(int) * C);
  L->idx = (int *)calloc((size_t)C * P, sizeof(int));
  L->ir = (int *)malloc(sizeof(int) * (size_t)C * E);
  L->syndLength = (int *)malloc(sizeof(int) * C);
  L->codeRate = (float *)malloc(sizeof(float) * C);
  if (read_ints(f, L->jc, N + 1)) { fclose(f); return -1; }
  for (int c = 0; c < C; c++) {
    if (fscanf(f, "%d", &L->bitsPerPeriod[c]) != 1) { fclose(f); return -1; }
    if (read_ints(f, L->idx + (size_t)c * P, L->bitsPerPeriod[c])) { fclose(f); return -1; }
    if (read_ints(f, L->ir + (size_t)c * E, E)) { fclose(f); return -1; }
    L->syndLength[c] = (N / P) * L->bitsPerPeriod[c];
    L->codeRate[c] = ((float)L->syndLength[c]) / ((float)N);
  }
  fclose(f);
  L->invWords = (N + 31) / 32;
  L->invH = (uint8_t *)malloc((size_t)4 * L->invWords * N);
  if (inv_path) {
    FILE *g = fopen(inv_path, "rb");
    if (!g) return -1;
    size_t want = (size_t)4 * L->invWords * N;
    size_t got = fread(L->invH, 1, want, g);
    fclose(g);
    if (got != want) return -1;
  }
  return 0;
}

/* ReadLadder_encoder (DvcEnc/src/ReadLdpcaLadderFile_LdpcaEncoder.cpp:12-36). */
int orc_read_encoder_ladder(orc_enc_ladder *L, const char *lad_path) {
  memset(L, 0, sizeof(*L));
  FILE *f = fopen(lad_path, "r");
  if (!f) return -1;
  if (fscanf(f, "%d %d", &L->N, &L->E) != 2) { fclose(f); return -1; }
  L->jc = (int *)malloc(sizeof(int) * (L->N + 1));
  L->ir = (int *)malloc(sizeof(int) * L->E);
  int rc = read_ints(f, L->jc, L->N + 1) || read_ints(f, L->ir, L->E);
  fclose(f);
  return rc ? -1 : 0;
}

/* Decoder row of encoder row r for code c: the smallest syndrome index k whose
 * transmitted accumulated-syndrome position tx_k = idx[k mod b] + (k/b)*period
 * is >= r (decoder row k merges encoder rows tx_{k-1}+1 .. tx_k). */
int orc_derive_decoder_ladder(orc_ladder *L, const orc_enc_ladder *enc, int numCodes, int period,
                              const int *bitsPerPeriod, const int *idx) {
  memset(L, 0, sizeof(*L));
  int N = enc->N, E = enc->E;
  L->numCodes = numCodes; L->N = N; L->E = E; L->period = period;
  L->jc = (int *)malloc(sizeof(int) * (N + 1));
  memcpy(L->jc, enc->jc, sizeof(int) * (N + 1));
  L->bitsPerPeriod = (int *)malloc(sizeof(int) * numCodes);
  L->idx = (int *)calloc((size_t)numCodes * period, sizeof(int));
  L->ir = (int *)malloc(sizeof(int) * (size_t)numCodes * E);
  L->syndLength = (int *)malloc(sizeof(int) * numCodes);
  L->codeRate = (float *)malloc(sizeof(float) * numCodes);
  int *rowmap = (int *)malloc(sizeof(int) * N);
  for (int c = 0; c < numCodes; c++) {
    int b = bitsPerPeriod[c];
    L->bitsPerPeriod[c] = b;
    memcpy(L->idx + (size_t)c * period, idx + (size_t)c * period, sizeof(int) * b);
    int m = (N / period) * b;
    L->syndLength[c] = m;
    L->codeRate[c] = ((float)m) / ((float)N);
    int k = 0;
    for (int r = 0; r < N; r++) {
      while (k < m && idx[(size_t)c * period + k % b] + (k / b) * period < r) k++;
      if (k >= m) { free(rowmap); return -1; }
      rowmap[r] = k;
    }
    for (int l = 0; l < E; l++) L->ir[(size_t)c * E + l] = rowmap[enc->ir[l]];
  }
  free(rowmap);
  L->invWords = (N + 31) / 32;
  L->invH = (uint8_t *)calloc((size_t)4 * L->invWords * N, 1);
  return 0;
}

/* Gauss-Jordan elimination of [H | I] over GF(2) with 64-bit bitsets; the
 * result row j is written MSB-first (bit i of the row is column i), padded to
 * words_per_row 32-bit words, as Inverse_Matrix_H_Reg1584.bin stores it. */
int orc_invert_h(const orc_enc_ladder *enc, uint8_t *out, int words_per_row) {
  int N = enc->N;
  int W = (N + 63) / 64;
  uint64_t *A = (uint64_t *)calloc((size_t)N * W, 8), *B = (uint64_t *)calloc((size_t)N * W, 8);
  for (int k = 0; k < N; k++)
    for (int l = enc->jc[k]; l < enc->jc[k + 1]; l++) {
      int r = enc->ir[l];
      A[(size_t)r * W + k / 64] ^= 1ull << (k % 64);
    }
  for (int r = 0; r < N; r++) B[(size_t)r * W + r / 64] |= 1ull << (r % 64);
  for (int col = 0; col < N; col++) {
    int piv = -1;
    for (int r = col; r < N; r++)
      if (A[(size_t)r * W + col / 64] >> (col % 64) & 1) { piv = r; break; }
    if (piv < 0) { free(A); free(B); return -1; }
    if (piv != col)
      for (int w = 0; w < W; w++) {
        uint64_t t = A[(size_t)piv * W + w]; A[(size_t)piv * W + w] = A[(size_t)col * W + w]; A[(size_t)col * W + w] = t;
        t = B[(size_t)piv * W + w]; B[(size_t)piv * W + w] = B[(size_t)col * W + w]; B[(size_t)col * W + w] = t;
      }
    const uint64_t *pa = A + (size_t)col * W, *pb = B + (size_t)col * W;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < N; r++) {
      if (r == col) continue;
      if (A[(size_t)r * W + col / 64] >> (col % 64) & 1) {
        uint64_t *ra = A + (size_t)r * W, *rb = B + (size_t)r * W;
        for (int w = 0; w < W; w++) { ra[w] ^= pa[w]; rb[w] ^= pb[w]; }
      }
    }
  }
  /* A is now I, B = H^-1 (row j: coefficients over syndrome indices i). */
  size_t rowBytes = (size_t)4 * words_per_row;
  memset(out, 0, rowBytes * N);
  for (int j = 0; j < N; j++)
    for (int i = 0; i < N; i++)
      if (B[(size_t)j * W + i / 64] >> (i % 64) & 1) out[rowBytes * j + i / 8] |= (uint8_t)(0x80 >> (i % 8));
  free(A); free(B);
  return 0;
}

void orc_free_ladder(orc_ladder *L) {
  free(L->jc); free(L->bitsPerPeriod); free(L->idx); free(L->ir);
  free(L->syndLength); free(L->codeRate); free(L->invH);
  memset(L, 0, sizeof(*L));
}
void orc_free_enc_ladder(orc_enc_ladder *L) { free(L->jc); free(L->ir); memset(L, 0, sizeof(*L)); }

/* BitToByte (LdpcaDecoder.cpp:41-70). */
int orc_bits_to_bytes(const uint8_t *bits, int nbits, uint8_t *bytes) {
  int nbytes = (nbits + 7) >> 3, bi = 0;
  uint8_t tmp = 0;
  for (int i = 0; i < nbits; i++) {
    tmp |= bits[i];
    if (((i + 1) & 7) == 0) { bytes[bi++] = tmp; tmp = 0; }
    else tmp <<= 1;
  }
  if (nbits & 7) bytes[bi++] = (uint8_t)(tmp << (7 - (nbits & 7)));
  return nbytes;
}

/* ByteToBit (LdpcaDecoder.cpp:81-107). */
void orc_bytes_to_bits(const uint8_t *bytes, int nbits, uint8_t *bits) {
  for (int i = 0; i < nbits; i++) bits[i] = (bytes[i >> 3] >> (7 - (i & 7))) & 1;
}

/* CRC8 (crc8.h:57-67) with the MSB-first table of x^8+x^2+x+1 (crc8.h:20-53). */
uint8_t orc_crc8(const uint8_t *buf, int len) {
  static uint8_t table[256];
  static int ready = 0;
  if (!ready) {
    for (int i = 0; i < 256; i++) {
      uint8_t c = (uint8_t)i;
      for (int b = 0; b < 8; b++) c = (uint8_t)((c & 0x80) ? (c << 1) ^ 0x07 : (c << 1));
      table[i] = c;
    }
    ready = 1;
  }
  uint8_t crc = 0;
  while (len--) crc = table[crc ^ *buf++];
  return crc;
}

#define ORC_LOG2 (0.693147181f)

/* ComputeCE (LdpcaDecoder.cpp:127-143); the C++ log(float) is logf. */
float orc_compute_ce(const float *prob, int N) {
  float rlb = 0.0f;
  for (int k = 0; k < N; k++) {
    float Pco = prob[k] / (1.0f + prob[k]);
    float tmp = logf(1 - Pco);
    rlb += (Pco * (-logf(Pco) + tmp) - tmp) / ORC_LOG2;
  }
  rlb /= (float)N;
  return rlb;
}

/* ComputeCE_OpenMP (LdpcaDecoder.cpp:109-126).  The partial sums are combined
 * in whatever order the OpenMP runtime uses. */
float orc_compute_ce_omp(const float *prob, int N, int nthreads) {
  float rlb = 0.0f;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for reduction(+ : rlb)
  for (int k = 0; k < N; k++) {
    float Pco = prob[k] / (1.0f + prob[k]);
    float tmp = logf(1 - Pco);
    rlb += (Pco * (-logf(Pco) + tmp) - tmp) / ORC_LOG2;
  }
  rlb /= (float)N;
  return rlb;
}

/* LdpcaDecoder.cpp:684-693. */
void orc_form_syndrome(const orc_ladder *L, int code, const uint8_t *acc, uint8_t *synd) {
  int m = L->syndLength[code], b = L->bitsPerPeriod[code];
  const int *idx = L->idx + (size_t)code * L->period;
  int cur = idx[0];
  synd[0] = acc[cur];
  for (int k = 1; k < m; k++) {
    int prev = cur;
    cur = idx[k % b] + (k / b) * L->period;
    synd[k] = acc[cur] ^ acc[prev];
  }
}

static inline int f32_sign(float v) { int32_t i; memcpy(&i, &v, 4); return (int)((i >> 31) | 1); }
static inline unsigned f32_signbit(float v) { uint32_t u; memcpy(&u, &v, 4); return u >> 31; }

/* One flooding iteration, steps 1-2 of beliefPropagation (:487-539):
 * from v2c (LLR_extrinsic) compute c2v and the new v2c / overall LLR. */
static void bp_iteration(const orc_ladder *L, const int *ir, int m, const float *LLR,
                         const uint8_t *synd, float *v2c, float *mag, signed char *sgn,
                         signed char *buf, float *rowTotal, float *overall, int omp) {
  int N = L->N, E = L->E;
  const int *jc = L->jc;
  for (int k = 0; k < E; k++) {
    sgn[k] = (signed char)f32_sign(v2c[k]);
    mag[k] = (float)sgn[k] * v2c[k];
  }
  for (int k = 0; k < m; k++) buf[k] = (signed char)(1 - (signed char)(synd[k] << 1));
  for (int k = 0; k < E; k++) buf[ir[k]] = (signed char)((buf[ir[k]] ^ sgn[k]) | 1);
  for (int k = 0; k < E; k++) sgn[k] = (signed char)((sgn[k] ^ buf[ir[k]]) | 1);
  for (int k = 0; k < m; k++) rowTotal[k] = 0;
#pragma omp parallel for if (omp)
  for (int k = 0; k < E; k++) mag[k] = -logf(tanhf((mag[k] + (float)0.000000001) / 2));
  for (int k = 0; k < E; k++) rowTotal[ir[k]] += mag[k];
#pragma omp parallel for if (omp)
  for (int k = 0; k < E; k++)
    mag[k] = -logf(tanhf((rowTotal[ir[k]] - mag[k] + (float)0.000000001) / 2)) * (float)sgn[k];
  for (int k = 0; k < N; k++) {
    float tmp = LLR[k];
    for (int l = jc[k]; l < jc[k + 1]; l++) tmp += mag[l];
    for (int l = jc[k]; l < jc[k + 1]; l++) v2c[l] = tmp - mag[l];
    overall[k] = tmp;
  }
}

/* Step 3 parity test (:551-568): 1 when every check is satisfied. */
static int syndrome_ok(const orc_ladder *L, const int *ir, int m, const uint8_t *synd,
                       const uint8_t *decoded, float *rowTotal) {
  for (int k = 0; k < m; k++) rowTotal[k] = synd[k];
  for (int k = 0; k < L->N; k++)
    for (int l = L->jc[k]; l < L->jc[k + 1]; l++) rowTotal[ir[l]] += (float)decoded[k];
  for (int k = 0; k < m; k++)
    if ((((int)rowTotal[k]) & 1) != 0) return 0;
  return 1;
}

typedef struct { float *v2c, *mag, *rowTotal, *overall; signed char *sgn, *buf; } bp_ws;
static void ws_alloc(bp_ws *w, const orc_ladder *L) {
  w->v2c = (float *)malloc(sizeof(float) * L->E);
  w->mag = (float *)malloc(sizeof(float) * L->E);
  w->rowTotal = (float *)malloc(sizeof(float) * L->N);
  w->overall = (float *)malloc(sizeof(float) * L->N);
  w->sgn = (signed char *)malloc(L->E);
  w->buf = (signed char *)malloc(L->N);
}
static void ws_free(bp_ws *w) { free(w->v2c); free(w->mag); free(w->rowTotal); free(w->overall); free(w->sgn); free(w->buf); }

int orc_bp_mode0(const orc_ladder *L, int code, const float *LLR, const uint8_t *synd,
                 uint8_t *decoded, int max_iter, int *iters) {
  const int *ir = L->ir + (size_t)code * L->E;
  int m = L->syndLength[code], N = L->N;
  bp_ws w; ws_alloc(&w, L);
  for (int k = 0; k < N; k++)
    for (int l = L->jc[k]; l < L->jc[k + 1]; l++) w.v2c[l] = LLR[k];
  int ok = 0, it;
  for (it = 0; it < max_iter; it++) {
    bp_iteration(L, ir, m, LLR, synd, w.v2c, w.mag, w.sgn, w.buf, w.rowTotal, w.overall, 0);
    for (int k = 0; k < N; k++) decoded[k] = (uint8_t)f32_signbit(w.overall[k]);
    if (syndrome_ok(L, ir, m, synd, decoded, w.rowTotal)) { ok = 1; it++; break; }
  }
  if (iters) *iters = it;
  ws_free(&w);
  return ok;
}

int orc_bp_mode1(const orc_ladder *L, int code, const float *LLR, const uint8_t *synd,
                 uint8_t *decoded, int max_iter, int nthreads, int *iters) {
  const int *ir = L->ir + (size_t)code * L->E;
  int m = L->syndLength[code], N = L->N;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  bp_ws w; ws_alloc(&w, L);
  for (int k = 0; k < N; k++)
    for (int l = L->jc[k]; l < L->jc[k + 1]; l++) w.v2c[l] = LLR[k];
  int sameCount = 0, it, result = 0, done = 0;
  for (it = 0; it < max_iter; it++) {
    bp_iteration(L, ir, m, LLR, synd, w.v2c, w.mag, w.sgn, w.buf, w.rowTotal, w.overall, 1);
    int same = 0;
    for (int k = 0; k < N; k++) {
      uint8_t d = (w.overall[k] < 0) ? 1 : 0;
      if (decoded[k] == d) same++;
      else decoded[k] = d;
    }
    sameCount = (same == N) ? sameCount + 1 : 0;
    if (sameCount == 4) { result = 0; done = 1; it++; break; }
    if (syndrome_ok(L, ir, m, synd, decoded, w.rowTotal)) { result = 1; done = 1; it++; break; }
  }
  if (!done) result = 0; /* :288-294: the final parity test repeats the failed one */
  if (iters) *iters = it;
  ws_free(&w);
  return result;
}

/* invH_mul (LdpcaDecoder.cpp:624-643): parity of (row j of H^-1) AND s. */
void orc_invh_mul(const orc_ladder *L, const uint8_t *synd, uint8_t *decoded) {
  int N = L->N, nw = L->invWords;
  uint8_t *sb = (uint8_t *)calloc((size_t)4 * nw, 1);
  orc_bits_to_bytes(synd, N, sb);
  uint32_t *sw = (uint32_t *)malloc((size_t)4 * nw);
  memcpy(sw, sb, (size_t)4 * nw);
  for (int j = 0; j < N; j++) {
    uint32_t t = 0, row[1];
    const uint8_t *rp = L->invH + (size_t)4 * nw * j;
    for (int k = 0; k < nw; k++) { memcpy(row, rp + 4 * k, 4); t ^= row[0] & sw[k]; }
    decoded[j] = (uint8_t)(__builtin_popcount(t) & 1);
  }
  free(sb); free(sw);
}

void orc_ldpca_decode(const orc_ladder *L, float *LR, const uint8_t *acc, uint8_t crc,
                      float entropy, int mode, int max_iter, int nthreads, uint8_t *decoded,
                      orc_ldpca_stats *st) {
  orc_ldpca_decode_steps(L, LR, acc, crc, entropy, mode, max_iter, nthreads, 0, decoded, st);
}

/* rate_step 1 (DYNAMICSTEP, LdpcaDecoder.cpp:19-21, 669-672, 760-768): after
 * the r-th failed request the code index grows by jump_step[min(r, 3)] =
 * {0,1,2,3} on top of the loop's increment, clamped so that the next code is
 * at most the full-rate one. */
void orc_ldpca_decode_steps(const orc_ladder *L, float *LR, const uint8_t *acc, uint8_t crc,
                            float entropy, int mode, int max_iter, int nthreads, int rate_step,
                            uint8_t *decoded, orc_ldpca_stats *st) {
  static const int jump_step[4] = {0, 1, 2, 3};
  int N = L->N;
  int request = 0, stepInc = 0;
  int start = 0;
  /* LdpcaDecoder.cpp:657-659 (bounded: the last code has rate 1). */
  while (start < L->numCodes - 1 && L->codeRate[start] < entropy) start++;
  for (int k = 0; k < N; k++) LR[k] = logf(LR[k]); /* :663-667 */
  uint8_t *synd = (uint8_t *)malloc(N);
  uint8_t *bytes = (uint8_t *)malloc((N + 7) / 8);
  memset(st, 0, sizeof(*st));
  st->start_level = start;
  for (int code = start; code < L->numCodes; code++) {
    int m = L->syndLength[code];
    st->rate = m + 8;
    st->level = code;
    st->requests++;
    orc_form_syndrome(L, code, acc, synd);
    if (m == N) { orc_invh_mul(L, synd, decoded); st->used_inverse = 1; break; }
    int it = 0, ok;
    if (mode == 0) ok = orc_bp_mode0(L, code, LR, synd, decoded, max_iter, &it);
    else ok = orc_bp_mode1(L, code, LR, synd, decoded, max_iter, nthreads, &it);
    st->total_iters += it;
    if (ok) {
      int nb = orc_bits_to_bytes(decoded, N, bytes);
      if (orc_crc8(bytes, nb) == crc) break;
    }
    if (rate_step) {
      if (request < 4) stepInc = jump_step[request];
      code += stepInc;
      if (code >= L->numCodes - 1) code = L->numCodes - 2;
      request++;
    }
  }
  free(synd); free(bytes);
}

/* LdpcaEncoder (DvcEnc/src/LdpcaEncoder.cpp:64-95). */
void orc_ldpca_encode(const orc_enc_ladder *L, const uint8_t *src, uint8_t *acc_bytes, uint8_t *crc) {
  int N = L->N;
  uint8_t *bytes = (uint8_t *)malloc((N + 7) / 8);
  int nb = orc_bits_to_bytes(src, N, bytes);
  *crc = orc_crc8(bytes, nb);
  free(bytes);
  uint8_t *acc = (uint8_t *)calloc(N, 1);
  for (int k = 0; k < N; k++)
    for (int l = L->jc[k]; l < L->jc[k + 1]; l++) acc[L->ir[l]] ^= src[k];
  for (int k = 1; k < N; k++) acc[k] ^= acc[k - 1];
  orc_bits_to_bytes(acc, N, acc_bytes);
  free(acc);
}
