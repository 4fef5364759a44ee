// ladder.cpp -- LDPCA ladder loading, decoder-code derivation, full-rate GF(2)
// inverse and the device image used by the BP kernels.
//
// The reference reads text ladders (ReadLdpcaLadderFile_LdpcaDecoder.cpp:131-250)
// and a packed inverse (Inverse_Matrix_H_Reg<N>.bin, :199-201).  This library
// rebuilds the same data from a compact pack (dvc_amd/data/<kind><N>.ladder: the
// encoder parity matrix and the per-code transmission pattern): decoder row k
// of code c merges the encoder rows (tx_{k-1}, tx_k], tx_k = idx[k mod b] +
// floor(k/b)*period, which is how the accumulated syndrome is sampled
// (LdpcaDecoder.cpp:684-693); the inverse is computed once by Gauss-Jordan
// elimination and cached next to the pack.
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <fstream>
#include "dvc_internal.h"

namespace dvc {

static bool read_file(const std::string &path, std::vector<uint8_t> &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return true;
}

static uint8_t crc8_bytes(const uint8_t *buf, int len) {
  static uint8_t table[256];
  static bool ready = false;
  if (!ready) {
    for (int i = 0; i < 256; i++) {
      uint8_t c = (uint8_t)i;
      for (int b = 0; b < 8; b++) c = (uint8_t)((c & 0x80) ? (c << 1) ^ 0x07 : (c << 1));
      table[i] = c;
    }
    ready = true;
  }
  uint8_t crc = 0;
  for (int i = 0; i < len; i++) crc = table[crc ^ buf[i]];
  return crc;
}

// H^-1 over GF(2) for the encoder matrix; row j packed MSB-first over the
// syndrome index.  Returns false if H is singular.
static bool invert_gf2(const Ladder &L, std::vector<uint8_t> &inv) {
  const int N = L.N, W = (N + 63) / 64, W2 = 2 * W;
  std::vector<uint64_t> M((size_t)N * W2, 0);
  for (int k = 0; k < N; k++)
    for (int l = L.jc[k]; l < L.jc[k + 1]; l++) M[(size_t)L.encIr[l] * W2 + k / 64] ^= 1ull << (k % 64);
  for (int r = 0; r < N; r++) M[(size_t)r * W2 + W + r / 64] |= 1ull << (r % 64);
  for (int col = 0; col < N; col++) {
    int piv = -1;
    for (int r = col; r < N; r++)
      if ((M[(size_t)r * W2 + col / 64] >> (col % 64)) & 1) { piv = r; break; }
    if (piv < 0) return false;
    if (piv != col)
      std::swap_ranges(M.begin() + (size_t)piv * W2, M.begin() + (size_t)(piv + 1) * W2, M.begin() + (size_t)col * W2);
    const uint64_t *p = &M[(size_t)col * W2];
    const int w0 = col / 64;  // words left of the pivot are already zero in the pivot row
    for (int r = 0; r < N; r++) {
      if (r == col) continue;
      uint64_t *q = &M[(size_t)r * W2];
      if ((q[w0] >> (col % 64)) & 1)
        for (int w = w0; w < W2; w++) q[w] ^= p[w];
    }
  }
  inv.assign((size_t)N * L.invRowBytes, 0);
  for (int j = 0; j < N; j++) {
    const uint64_t *b = &M[(size_t)j * W2 + W];
    uint8_t *o = &inv[(size_t)j * L.invRowBytes];
    for (int i = 0; i < N; i++)
      if ((b[i / 64] >> (i % 64)) & 1) o[i / 8] |= (uint8_t)(0x80 >> (i % 8));
  }
  return true;
}

int load_ladder(const std::string &dir, const char *kind, int N, Ladder &L) {
  const std::string base = dir + "/" + kind + std::to_string(N);
  const std::string path = base + ".ladder";
  std::vector<uint8_t> blob;
  DVC_CHECK(read_file(path, blob), DVC_EIO, "cannot read ladder pack " + path);
  DVC_CHECK(blob.size() >= 24 && memcmp(blob.data(), "DVCL", 4) == 0, DVC_EIO, "bad ladder pack " + path);
  uint32_t hdr[5];
  memcpy(hdr, blob.data() + 4, sizeof(hdr));
  DVC_CHECK(hdr[0] == 1 && (int)hdr[1] == N, DVC_EIO, "ladder pack version/size mismatch " + path);
  L.N = N; L.E = (int)hdr[2]; L.period = (int)hdr[3]; L.numCodes = (int)hdr[4];
  size_t need = 24 + 4 * (size_t)(N + 1) + 2 * (size_t)L.E + 2 * (size_t)L.numCodes + 2 * (size_t)L.numCodes * L.period;
  DVC_CHECK(blob.size() == need, DVC_EIO, "truncated ladder pack " + path);
  const uint8_t *p = blob.data() + 24;
  L.jc.resize(N + 1);
  for (int i = 0; i <= N; i++) { uint32_t v; memcpy(&v, p, 4); p += 4; L.jc[i] = (int)v; }
  L.encIr.resize(L.E);
  for (int i = 0; i < L.E; i++) { uint16_t v; memcpy(&v, p, 2); p += 2; L.encIr[i] = v; }
  L.bitsPerPeriod.resize(L.numCodes);
  for (int i = 0; i < L.numCodes; i++) { uint16_t v; memcpy(&v, p, 2); p += 2; L.bitsPerPeriod[i] = v; }
  L.idx.resize((size_t)L.numCodes * L.period);
  for (size_t i = 0; i < L.idx.size(); i++) { uint16_t v; memcpy(&v, p, 2); p += 2; L.idx[i] = v; }
  L.regular = true;
  for (int k = 0; k < N; k++) {
    DVC_CHECK(L.jc[k + 1] > L.jc[k], DVC_EIO, "ladder pack: empty column");
    if (L.jc[k + 1] - L.jc[k] != 3) L.regular = false;
  }
  DVC_CHECK(L.jc[0] == 0 && L.jc[N] == L.E && N % L.period == 0, DVC_EIO, "ladder pack: inconsistent sizes");

  // Decoder codes (codeRate as ReadLdpcaLadderFile_LdpcaDecoder.cpp:182-183).
  L.syndLength.resize(L.numCodes);
  L.codeRate.resize(L.numCodes);
  L.decIr.resize((size_t)L.numCodes * L.E);
  std::vector<int> rowmap(N);
  for (int c = 0; c < L.numCodes; c++) {
    int b = L.bitsPerPeriod[c], m = (N / L.period) * b;
    L.syndLength[c] = m;
    L.codeRate[c] = ((float)m) / ((float)N);
    const int *ix = &L.idx[(size_t)c * L.period];
    int k = 0;
    for (int r = 0; r < N; r++) {
      while (k < m && ix[k % b] + (k / b) * L.period < r) k++;
      DVC_CHECK(k < m, DVC_EIO, "ladder pack: transmission pattern does not cover the code");
      rowmap[r] = k;
    }
    for (int l = 0; l < L.E; l++) L.decIr[(size_t)c * L.E + l] = rowmap[L.encIr[l]];
  }
  DVC_CHECK(L.syndLength[L.numCodes - 1] == N, DVC_EIO, "ladder pack: last code is not full rate");

  // Full-rate inverse, cached as <pack>.invh.
  L.invRowBytes = 4 * ((N + 31) / 32);
  const std::string invPath = base + ".invh";
  std::vector<uint8_t> inv;
  if (!(read_file(invPath, inv) && inv.size() == (size_t)N * L.invRowBytes)) {
    DVC_CHECK(invert_gf2(L, inv), DVC_EIO, "ladder pack: parity matrix is singular");
    std::ofstream f(invPath, std::ios::binary);
    if (f) f.write((const char *)inv.data(), (std::streamsize)inv.size());
  }
  L.invH.swap(inv);

  // CRC-8 is linear with zero initial value and final xor (crc8.h:5-6), so the
  // CRC of a bitplane is the XOR of the CRCs of its set bits.
  L.crcBasis.resize(N);
  std::vector<uint8_t> msg(N / 8, 0);
  for (int k = 0; k < N; k++) {
    msg[k >> 3] = (uint8_t)(0x80 >> (k & 7));
    L.crcBasis[k] = crc8_bytes(msg.data(), N / 8);
    msg[k >> 3] = 0;
  }
  return DVC_OK;
}

// Deals the columns of an irregular ladder to kIrregThreads threads: by
// decreasing degree, each to the thread with the fewest edges so far, so the
// threads of a workgroup carry nearly equal variable-node work.
static void column_slots(const Ladder &L, std::vector<int16_t> &slots, int &perThread) {
  const int T = kIrregThreads, N = L.N;
  std::vector<int> order(N), load(T, 0), cnt(T, 0), owner(N);
  for (int k = 0; k < N; k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(),
                   [&L](int a, int b) { return L.jc[a + 1] - L.jc[a] > L.jc[b + 1] - L.jc[b]; });
  for (int k : order) {
    int t = 0;
    for (int u = 1; u < T; u++)
      if (load[u] < load[t]) t = u;
    owner[k] = t;
    load[t] += L.jc[k + 1] - L.jc[k];
    cnt[t]++;
  }
  perThread = *std::max_element(cnt.begin(), cnt.end());
  slots.assign((size_t)T * perThread, (int16_t)-1);
  std::fill(cnt.begin(), cnt.end(), 0);
  for (int k = 0; k < N; k++) slots[(size_t)owner[k] * perThread + cnt[owner[k]]++] = (int16_t)k;
}

int upload_ladder(Ladder &L) {
  const int N = L.N, E = L.E, C = L.numCodes;
  std::vector<uint32_t> colRows(L.regular ? (size_t)C * N * 3 : 0), colEdges(L.regular ? 0 : (size_t)C * E);
  std::vector<uint16_t> rowPtr((size_t)C * (N + 1)), tx((size_t)C * N);
  DVC_CHECK(E <= 0xffff && N <= 0x2000, DVC_EIO, "ladder: more than 65535 edges or 8192 checks");
  for (int c = 0; c < C; c++) {
    int m = L.syndLength[c], b = L.bitsPerPeriod[c];
    const int *dir = &L.decIr[(size_t)c * E];
    std::vector<int> cnt(m + 1, 0);
    for (int e = 0; e < E; e++) cnt[dir[e] + 1]++;
    for (int r = 0; r < m; r++) cnt[r + 1] += cnt[r];
    for (int r = 0; r <= m; r++) rowPtr[(size_t)c * (N + 1) + r] = (uint16_t)cnt[r];
    // rank of each edge within its row, rows listing edges in ascending order
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    std::vector<uint32_t> field(E);
    for (int e = 0; e < E; e++) {
      const int rank = fill[dir[e]]++ - cnt[dir[e]];
      if (!L.regular) {
        colEdges[(size_t)c * E + e] = (uint32_t)dir[e] | ((uint32_t)rank << 16);
        continue;
      }
      field[e] = (uint32_t)dir[e] | ((uint32_t)(cnt[dir[e]] + rank) << 15);  // row | message slot << 15
    }
    for (int k = 0; L.regular && k < N; k++)  // every column has degree 3: edges 3k..3k+2
      for (int q = 0; q < 3; q++) colRows[((size_t)c * N + k) * 3 + q] = field[3 * k + q];
    const int *ix = &L.idx[(size_t)c * L.period];
    for (int k = 0; k < m; k++) tx[(size_t)c * N + k] = (uint16_t)(ix[k % b] + (k / b) * L.period);
  }
  if (L.regular) {
    DVC_HIP(hipMalloc(&L.d_colRows, colRows.size() * 4));
    } else {
    std::vector<int16_t> slots;
    column_slots(L, slots, L.colSlotsPerThread);
    DVC_HIP(hipMalloc(&L.d_colEdges, colEdges.size() * 4));
    DVC_HIP(hipMalloc(&L.d_jc, (N + 1) * sizeof(int)));
    DVC_HIP(hipMalloc(&L.d_colSlots, slots.size() * sizeof(int16_t)));
    DVC_HIP(hipMemcpy(L.d_colEdges, colEdges.data(), colEdges.size() * 4, hipMemcpyHostToDevice));
    DVC_HIP(hipMemcpy(L.d_jc, L.jc.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice));
    DVC_HIP(hipMemcpy(L.d_colSlots, slots.data(), slots.size() * sizeof(int16_t), hipMemcpyHostToDevice));
  }
  DVC_HIP(hipMalloc(&L.d_rowPtr, rowPtr.size() * 2));
  DVC_HIP(hipMalloc(&L.d_tx, tx.size() * 2));
  DVC_HIP(hipMalloc(&L.d_syndLength, C * sizeof(int)));
  DVC_HIP(hipMalloc(&L.d_codeRate, C * sizeof(float)));
  DVC_HIP(hipMalloc(&L.d_invH, L.invH.size()));
  DVC_HIP(hipMalloc(&L.d_crcBasis, N));
  DVC_HIP(hipMemcpy(L.d_colRows, colRows.data(), colRows.size() * 4, hipMemcpyHostToDevice));
  DVC_HIP(hipMemcpy(L.d_rowPtr, rowPtr.data(), rowPtr.size() * 2, hipMemcpyHostToDevice));
  DVC_HIP(hipMemcpy(L.d_tx, tx.data(), tx.size() * 2, hipMemcpyHostToDevice));
  DVC_HIP(hipMemcpy(L.d_syndLength, L.syndLength.data(), C * sizeof(int), hipMemcpyHostToDevice));
  DVC_HIP(hipMemcpy(L.d_codeRate, L.codeRate.data(), C * sizeof(float), hipMemcpyHostToDevice));
  DVC_HIP(hipMemcpy(L.d_invH, L.invH.data(), L.invH.size(), hipMemcpyHostToDevice));
  DVC_HIP(hipMemcpy(L.d_crcBasis, L.crcBasis.data(), N, hipMemcpyHostToDevice));
  return DVC_OK;
}

void free_ladder(Ladder &L) {
  void *ptrs[] = {L.d_colRows, L.d_colEdges, L.d_jc, L.d_colSlots, L.d_rowPtr,  L.d_tx,
                  L.d_syndLength, L.d_codeRate, L.d_invH, L.d_crcBasis};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  L.d_colRows = nullptr;
  L.d_colEdges = nullptr; L.d_jc = nullptr; L.d_colSlots = nullptr;
  L.d_rowPtr = L.d_tx = nullptr;
  L.d_syndLength = nullptr; L.d_codeRate = nullptr; L.d_invH = L.d_crcBasis = nullptr;
}

}  // namespace dvc
