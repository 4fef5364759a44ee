// autotune.cpp -- chooses the BP launch configuration of a context by timing
// the variants on a calibration batch: threads per workgroup for each code
// length, QCIF codewords resident per CU, and the interval at which running
// speculative items poll for cancellation.  Every variant decodes the same
// bits (the tuner checks that it does); only the time differs.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <chrono>
#include <vector>
#include "dvc_internal.h"
#include "ldpca_kernels.h"

using namespace dvc;

namespace {

struct Lcg {
  uint64_t s;
  double uniform() {  // [0, 1)
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  }
};

// One bitplane of the calibration batch: Bernoulli(p) bits, Laplacian side
// information y = b*4 + Laplace(1/alpha), the likelihood ratio the decoder
// takes (clipped as condBitProb clips it), the accumulated syndrome through
// the encoder's code (MSB-first bytes) and the CRC-8.
void make_plane(const Ladder &L, Lcg &g, std::vector<float> &lr, std::vector<uint8_t> &acc, uint8_t &crc,
                std::vector<uint8_t> &bits) {
  const int N = L.N;
  const double p = 0.02 + 0.28 * g.uniform(), alpha = 0.1 + 1.4 * g.uniform(), delta = 4.0;
  std::vector<uint8_t> synd(N, 0);
  bits.assign(N, 0);
  lr.resize(N);
  const float lo = (float)(0.0001 / 0.9999), hi = (float)(0.9999 / 0.0001);
  crc = 0;
  for (int k = 0; k < N; k++) {
    bits[k] = g.uniform() < p;
    const double u = g.uniform() - 0.5;
    const double lap = (u < 0 ? 1.0 : -1.0) * log(1.0 - 2.0 * fabs(u)) / alpha;
    const double y = bits[k] * delta + lap;
    double llr = alpha * (fabs(y - delta) - fabs(y)) + log((1.0 - p) / p);
    llr = std::min(30.0, std::max(-30.0, llr));
    lr[k] = std::min(hi, std::max(lo, (float)exp(llr)));
    if (bits[k]) {
      crc ^= L.crcBasis[k];
      for (int l = L.jc[k]; l < L.jc[k + 1]; l++) synd[L.encIr[l]] ^= 1;
    }
  }
  acc.assign(N / 8, 0);
  uint8_t run = 0;
  for (int k = 0; k < N; k++) {
    run ^= synd[k];
    if (run) acc[k >> 3] |= (uint8_t)(0x80 >> (k & 7));
  }
}

struct Batch {
  std::vector<std::vector<float>> lr;
  std::vector<std::vector<uint8_t>> acc, out, bits;
  std::vector<uint8_t> crc;
  std::vector<dvc_ldpca_item> items;
};

void make_batch(const Ladder &L, int n, uint64_t seed, Batch &b) {
  Lcg g{seed};
  b.lr.resize(n); b.acc.resize(n); b.bits.resize(n); b.out.assign(n, std::vector<uint8_t>(L.N)); b.crc.resize(n);
  b.items.resize(n);
  for (int i = 0; i < n; i++) {
    make_plane(L, g, b.lr[i], b.acc[i], b.crc[i], b.bits[i]);
    b.items[i] = dvc_ldpca_item{L.N, b.lr[i].data(), b.acc[i].data(), b.crc[i], -1, b.out[i].data(), {}};
  }
}

// Best-of-two wall time of the batched decode of `b` with the context's
// current tuning, in ms; the decoded planes are left in b.out.
int time_batch(dvc_ctx *c, Batch &b, double *ms) {
  double best = 1e30;
  for (int r = 0; r < 2; r++) {
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = dvc_ldpca_decode_batch(c, (int)b.items.size(), b.items.data());
    if (rc) return rc;
    const double dt = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    best = std::min(best, dt);
  }
  *ms = best;
  return DVC_OK;
}

}  // namespace

extern "C" {

int dvc_get_tuning(dvc_ctx *c, dvc_tuning *t) {
  DVC_CHECK(c && t, DVC_EINVAL, "dvc_get_tuning: NULL argument");
  memset(t, 0, sizeof(*t));
  t->cif_threads = c->bp_threads_cif;
  t->qcif_threads = c->bp_threads_qcif;
  t->qcif_wg_per_cu = c->bp_wg_per_cu_qcif;
  t->poll = c->bp_poll;
  return DVC_OK;
}

int dvc_set_tuning(dvc_ctx *c, const dvc_tuning *t) {
  DVC_CHECK(c && t, DVC_EINVAL, "dvc_set_tuning: NULL argument");
  DVC_CHECK(t->cif_threads == 512 || t->cif_threads == 704 || t->cif_threads == 1024, DVC_EINVAL,
            "dvc_set_tuning: cif_threads must be 512, 704 or 1024");
  DVC_CHECK(t->qcif_threads == 256 || t->qcif_threads == 512, DVC_EINVAL,
            "dvc_set_tuning: qcif_threads must be 256 or 512");
  DVC_CHECK(t->qcif_wg_per_cu >= 1 && t->qcif_wg_per_cu <= 4, DVC_EINVAL, "dvc_set_tuning: qcif_wg_per_cu must be 1..4");
  DVC_CHECK(t->poll >= 1 && t->poll <= 8, DVC_EINVAL, "dvc_set_tuning: poll must be 1..8");
  c->bp_threads_cif = t->cif_threads;
  c->bp_threads_qcif = t->qcif_threads;
  c->bp_wg_per_cu_qcif = t->qcif_wg_per_cu;
  c->bp_poll = t->poll;
  return DVC_OK;
}

int dvc_autotune(dvc_ctx *c, int planes, dvc_tuning *out) {
  DVC_CHECK(c, DVC_EINVAL, "dvc_autotune: NULL context");
  const int n = planes > 0 ? planes : 128;
  dvc_tuning cur;
  dvc_get_tuning(c, &cur);
  Batch cif, qcif;
  make_batch(c->lad6336, n, 0x6336ull, cif);
  make_batch(c->lad1584, n, 0x1584ull, qcif);
  int rc;
  // every variant must decode exactly what the first one decodes
  auto check_same = [](const Batch &b, const std::vector<std::vector<uint8_t>> &ref) { return b.out == ref; };
  std::vector<std::vector<uint8_t>> refC, refQ;
  double bestC = 1e30, bestQ = 1e30;
  dvc_tuning best = cur;
  for (int T : {1024, 704, 512}) {
    dvc_tuning t = cur;
    t.cif_threads = T;
    dvc_set_tuning(c, &t);
    double ms;
    if ((rc = time_batch(c, cif, &ms))) { dvc_set_tuning(c, &cur); return rc; }
    if (refC.empty()) {
      refC = cif.out;
      if (refC != cif.bits) { dvc_set_tuning(c, &cur); set_error("dvc_autotune: calibration batch decoded wrongly"); return DVC_ESTATE; }
    }
    else if (!check_same(cif, refC)) { dvc_set_tuning(c, &cur); set_error("dvc_autotune: variants decoded differently"); return DVC_ESTATE; }
    if (ms < bestC) { bestC = ms; best.cif_threads = T; }
  }
  for (int T : {256, 512})
    for (int w : {4, 2}) {
      dvc_tuning t = cur;
      t.qcif_threads = T;
      t.qcif_wg_per_cu = w;
      dvc_set_tuning(c, &t);
      double ms;
      if ((rc = time_batch(c, qcif, &ms))) { dvc_set_tuning(c, &cur); return rc; }
      if (refQ.empty()) {
        refQ = qcif.out;
        if (refQ != qcif.bits) { dvc_set_tuning(c, &cur); set_error("dvc_autotune: calibration batch decoded wrongly"); return DVC_ESTATE; }
      }
      else if (!check_same(qcif, refQ)) { dvc_set_tuning(c, &cur); set_error("dvc_autotune: variants decoded differently"); return DVC_ESTATE; }
      if (ms < bestQ) { bestQ = ms; best.qcif_threads = T; best.qcif_wg_per_cu = w; }
    }
  double bestTot = 1e30;
  for (int poll : {1, 2, 4}) {
    dvc_tuning t = best;
    t.poll = poll;
    dvc_set_tuning(c, &t);
    double a, b;
    if ((rc = time_batch(c, cif, &a)) || (rc = time_batch(c, qcif, &b))) { dvc_set_tuning(c, &cur); return rc; }
    if (!check_same(cif, refC) || !check_same(qcif, refQ)) {
      dvc_set_tuning(c, &cur);
      set_error("dvc_autotune: variants decoded differently");
      return DVC_ESTATE;
    }
    if (a + b < bestTot) { bestTot = a + b; best.poll = poll; best.cif_ms = a; best.qcif_ms = b; }
  }
  dvc_set_tuning(c, &best);
  if (out) *out = best;
  return DVC_OK;
}

}  // extern "C"
