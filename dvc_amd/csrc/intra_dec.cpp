// intra_dec.cpp -- the intra blocks' bitstream: CAVLC parsing and
// dequantisation of the 4x4 luma blocks a WZ frame codes in intra mode
// (intraRecon's read_one_block, intra.cpp:185-215, over readCoeff4x4_CAVLC,
// vlc.cpp:613-761).  Host code: the intra bitstream is one variable-length
// bit string read in the decoder's frame order, each block's code tables
// chosen from the coefficient counts of the blocks parsed before it.  The
// prediction, inverse transform and reconstruction of the parsed blocks run
// on the device (k_intra_recon).
#include <cmath>
#include <cstring>
#include <vector>
#include "dvc_internal.h"

namespace dvc {
namespace {

// CAVLC code tables of ITU-T H.264 (Table 9-5 coeff_token for 0 <= nC < 2,
// 2 <= nC < 4 and 4 <= nC < 8; Table 9-7 total_zeros of 4x4 blocks; Table
// 9-10 run_before), written as the codeword bit strings; "" = no codeword.
// coeff_token: [nC class][TotalCoeff 0..16][TrailingOnes 0..3]
const char *const kCoeffToken[3][17][4] = {
    {{"1", "", "", ""},
     {"000101", "01", "", ""},
     {"00000111", "000100", "001", ""},
     {"000000111", "00000110", "0000101", "00011"},
     {"0000000111", "000000110", "00000101", "000011"},
     {"00000000111", "0000000110", "000000101", "0000100"},
     {"0000000001111", "00000000110", "0000000101", "00000100"},
     {"0000000001011", "0000000001110", "00000000101", "000000100"},
     {"0000000001000", "0000000001010", "0000000001101", "0000000100"},
     {"00000000001111", "00000000001110", "0000000001001", "00000000100"},
     {"00000000001011", "00000000001010", "00000000001101", "0000000001100"},
     {"000000000001111", "000000000001110", "00000000001001", "00000000001100"},
     {"000000000001011", "000000000001010", "000000000001101", "00000000001000"},
     {"0000000000001111", "000000000000001", "000000000001001", "000000000001100"},
     {"0000000000001011", "0000000000001110", "0000000000001101", "000000000001000"},
     {"0000000000000111", "0000000000001010", "0000000000001001", "0000000000001100"},
     {"0000000000000100", "0000000000000110", "0000000000000101", "0000000000001000"}},
    {{"11", "", "", ""},
     {"001011", "10", "", ""},
     {"000111", "00111", "011", ""},
     {"0000111", "001010", "001001", "0101"},
     {"00000111", "000110", "000101", "0100"},
     {"00000100", "0000110", "0000101", "00110"},
     {"000000111", "00000110", "00000101", "001000"},
     {"00000001111", "000000110", "000000101", "000100"},
     {"00000001011", "00000001110", "00000001101", "0000100"},
     {"000000001111", "00000001010", "00000001001", "000000100"},
     {"000000001011", "000000001110", "000000001101", "00000001100"},
     {"000000001000", "000000001010", "000000001001", "00000001000"},
     {"0000000001111", "0000000001110", "0000000001101", "000000001100"},
     {"0000000001011", "0000000001010", "0000000001001", "0000000001100"},
     {"0000000000111", "00000000001011", "0000000000110", "0000000001000"},
     {"00000000001001", "00000000001000", "00000000001010", "0000000000001"},
     {"00000000000111", "00000000000110", "00000000000101", "00000000000100"}},
    {{"1111", "", "", ""},
     {"001111", "1110", "", ""},
     {"001011", "01111", "1101", ""},
     {"001000", "01100", "01110", "1100"},
     {"0001111", "01010", "01011", "1011"},
     {"0001011", "01000", "01001", "1010"},
     {"0001001", "001110", "001101", "1001"},
     {"0001000", "001010", "001001", "1000"},
     {"00001111", "0001110", "0001101", "01101"},
     {"00001011", "00001110", "0001010", "001100"},
     {"000001111", "00001010", "00001101", "0001100"},
     {"000001011", "000001110", "00001001", "00001100"},
     {"000001000", "000001010", "000001101", "00001000"},
     {"0000001101", "000000111", "000001001", "000001100"},
     {"0000001001", "0000001100", "0000001011", "0000001010"},
     {"0000000101", "0000001000", "0000000111", "0000000110"},
     {"0000000001", "0000000100", "0000000011", "0000000010"}}
};
// total_zeros: [TotalCoeff - 1][total_zeros]
const char *const kTotalZeros[15][16] = {
    {"1", "011", "010", "0011", "0010", "00011", "00010", "000011", "000010", "0000011", "0000010", "00000011", "00000010", "000000011", "000000010", "000000001"},
    {"111", "110", "101", "100", "011", "0101", "0100", "0011", "0010", "00011", "00010", "000011", "000010", "000001", "000000"},
    {"0101", "111", "110", "101", "0100", "0011", "100", "011", "0010", "00011", "00010", "000001", "00001", "000000"},
    {"00011", "111", "0101", "0100", "110", "101", "100", "0011", "011", "0010", "00010", "00001", "00000"},
    {"0101", "0100", "0011", "111", "110", "101", "100", "011", "0010", "00001", "0001", "00000"},
    {"000001", "00001", "111", "110", "101", "100", "011", "010", "0001", "001", "000000"},
    {"000001", "00001", "101", "100", "011", "11", "010", "0001", "001", "000000"},
    {"000001", "0001", "00001", "011", "11", "10", "010", "001", "000000"},
    {"000001", "000000", "0001", "11", "10", "001", "01", "00001"},
    {"00001", "00000", "001", "11", "10", "01", "0001"},
    {"0000", "0001", "001", "010", "1", "011"},
    {"0000", "0001", "01", "1", "001"},
    {"000", "001", "1", "01"},
    {"00", "01", "1"},
    {"0", "1"}
};
// run_before: [min(zerosLeft, 7) - 1][run_before]
const char *const kRunBefore[7][15] = {
    {"1", "0"},
    {"1", "01", "00"},
    {"11", "10", "01", "00"},
    {"11", "10", "01", "001", "000"},
    {"11", "10", "011", "010", "001", "000"},
    {"11", "000", "001", "011", "010", "101", "100"},
    {"111", "110", "101", "100", "011", "010", "001", "0001", "00001", "000001", "0000001", "00000001", "000000001", "0000000001", "00000000001"}
};

// MSB-first bit reader over the intra bitstream; bits past its end read 0.
struct Bits {
  const uint8_t *d;
  size_t nbits;
  size_t pos;
  int bit(size_t p) const { return p < nbits ? (d[p >> 3] >> (7 - (p & 7))) & 1 : 0; }
  uint32_t peek(int n) const {
    uint32_t v = 0;
    for (int k = 0; k < n; k++) v = (v << 1) | (uint32_t)bit(pos + k);
    return v;
  }
  uint32_t read(int n) {
    const uint32_t v = peek(n);
    pos += n;
    return v;
  }
  bool matches(const char *code) const {
    for (int k = 0; code[k]; k++)
      if (bit(pos + k) != code[k] - '0') return false;
    return true;
  }
};

// The index of the codeword of `table` (n entries) that the stream holds next.
int read_code(Bits &b, const char *const *table, int n) {
  for (int i = 0; i < n; i++)
    if (table[i][0] && b.matches(table[i])) {
      b.pos += strlen(table[i]);
      return i;
    }
  return -1;
}

// level_prefix / level_suffix decoding with suffixLength 0
// (readSyntaxElement_Level_VLC0, vlc.cpp:301-364).
int level_vlc0(Bits &b) {
  int len = 1;
  while (!b.read(1)) len++;  // len = leading zeros + 1
  int level, sign;
  if (len < 15) {
    sign = (len - 1) & 1;
    level = ((len - 1) >> 1) + 1;
  } else if (len == 15) {
    const int code = (1 << 4) | (int)b.read(4);
    sign = code & 1;
    level = ((code >> 1) & 7) + 8;
  } else {
    const int addbit = len - 16, n = len - 4;
    const int code = (int)b.read(n);
    sign = code & 1;
    level = (code >> 1) + (2048 << addbit) - 2032;
  }
  return sign ? -level : level;
}

// suffixLength vlc >= 1 (readSyntaxElement_Level_VLCN, vlc.cpp:231-299).
int level_vlcn(Bits &b, int vlc) {
  int len = 1;
  while (!b.read(1)) len++;
  const int shift = vlc - 1;
  int levabs;
  if (len < 16) {
    levabs = ((len - 1) << shift) + 1;
    if (shift) levabs += (int)b.read(shift);
  } else {
    const int addbit = len - 5;
    levabs = (int)b.read(addbit) + (1 << addbit) + (15 << shift) - 2047;
  }
  return b.read(1) ? -levabs : levabs;
}

constexpr uint8_t kIntraB = 1;  // BlockModeFlag of an intra block (defines.h:16)

}  // namespace

// Parses the CAVLC blocks of one WZ frame's intra blocks (raster order of
// the 4x4 blocks whose 8x8 block is intra) from bit `*bitpos` of the intra
// bitstream, advancing it, and writes each block's dequantised coefficients
// -- ROUND((double)level * Qstep * preScale), Qstep = 2^((QP - 4) / 6)
// (intra.cpp:185-215, 240-258) -- to coeffs[16 * block] (row-major 4x4;
// zero for other blocks).
int intra_parse_frame(const uint8_t *data, size_t len, size_t *bitpos, const uint8_t *mode, int W, int H, int qp,
                      int32_t *coeffs) {
  const int bw = W / 4, bh = H / 4, w8 = W / 8;
  Bits b{data, len * 8, *bitpos};
  std::vector<int> nonzero((size_t)bw * bh, 0);
  // initTransform's preScale (transform.cpp:12-52) and the intra step
  const double a = 0.5, bb = std::sqrt(0.4), a2 = 0.25, b2 = 0.4, ab = a * bb;
  const double pre[16] = {a2 * 64.0, ab * 64.0, a2 * 64.0, ab * 64.0, ab * 64.0, b2 * 64.0, ab * 64.0, b2 * 64.0,
                          a2 * 64.0, ab * 64.0, a2 * 64.0, ab * 64.0, ab * 64.0, b2 * 64.0, ab * 64.0, b2 * 64.0};
  const double qstep = std::pow(2.0, ((double)qp - 4.0) / 6.0);
  memset(coeffs, 0, sizeof(int32_t) * 16 * (size_t)bw * bh);
  auto intra = [&](int x, int y) { return mode[(y >> 1) * w8 + (x >> 1)] == kIntraB; };
  for (int y = 0; y < bh; y++)
    for (int x = 0; x < bw; x++) {
      if (!intra(x, y)) continue;
      // nC from the left and upper intra blocks' coefficient counts (predict_nnz, vlc.cpp:564-611)
      int nc = 0, cnt = 0;
      if (x > 0 && intra(x - 1, y)) { nc += nonzero[(size_t)y * bw + x - 1]; cnt++; }
      if (y > 0 && intra(x, y - 1)) { nc += nonzero[(size_t)(y - 1) * bw + x]; cnt++; }
      if (cnt == 2) nc = (nc + 1) >> 1;
      int total = 0, ones = 0;
      if (nc >= 8) {  // 6-bit fixed-length coeff_token
        const int code = (int)b.read(6);
        ones = code & 3;
        total = code >> 2;
        if (total == 0 && ones == 3) ones = 0;
        else total++;
      } else {
        const int cls = nc < 2 ? 0 : (nc < 4 ? 1 : 2);
        int found = -1;
        for (int tc = 0; tc <= 16 && found < 0; tc++)
          for (int t1 = 0; t1 < 4; t1++) {
            const char *cw = kCoeffToken[cls][tc][t1];
            if (cw[0] && b.matches(cw)) {
              b.pos += strlen(cw);
              total = tc;
              ones = t1;
              found = 1;
              break;
            }
          }
        DVC_CHECK(found > 0, DVC_EINVAL, "intra bitstream: no coeff_token codeword");
      }
      nonzero[(size_t)y * bw + x] = total;
      int level[16] = {0}, run[16] = {0};
      if (total) {
        // trailing +-1s, the last coefficients first
        if (ones) {
          const uint32_t signs = b.read(ones);
          for (int k = 0; k < ones; k++) level[total - 1 - k] = (signs >> (ones - 1 - k)) & 1 ? -1 : 1;
        }
        bool bump = !(total > 3 && ones == 3);  // the first level after fewer than 3 trailing ones is >= 2
        int vlc = (total > 10 && ones < 3) ? 1 : 0;
        static const int kIncVlc[7] = {0, 3, 6, 12, 24, 48, 32768};
        for (int k = total - 1 - ones; k >= 0; k--) {
          int lv = vlc == 0 ? level_vlc0(b) : level_vlcn(b, vlc);
          if (bump) {
            lv += lv > 0 ? 1 : -1;
            bump = false;
          }
          level[k] = lv;
          const int absl = lv < 0 ? -lv : lv;
          if (absl > kIncVlc[vlc]) vlc++;
          if (k == total - 1 - ones && absl > 3) vlc = 2;
        }
        int zeros = 0;
        if (total < 16) {
          zeros = read_code(b, kTotalZeros[total - 1], 16 - (total - 1));
          DVC_CHECK(zeros >= 0, DVC_EINVAL, "intra bitstream: no total_zeros codeword");
        }
        int left = zeros, i = total - 1;
        while (left > 0 && i > 0) {
          const int t = left - 1 < 6 ? left - 1 : 6;
          const int r = read_code(b, kRunBefore[t], t < 6 ? t + 2 : 15);
          DVC_CHECK(r >= 0, DVC_EINVAL, "intra bitstream: no run_before codeword");
          run[i] = r;
          left -= r;
          i--;
        }
        run[i] = left;
      }
      // coefficients in zig-zag order from the levels and runs (lowest frequency first)
      int coef[16] = {0}, pos = 0;
      for (int k = 0; k < total; k++)
        if (level[k]) {
          pos += run[k];
          DVC_CHECK(pos < 16, DVC_EINVAL, "intra bitstream: coefficient runs beyond the block");
          coef[zigzag_entry(pos)] = level[k];
          pos++;
        }
      int32_t *out = coeffs + 16 * ((size_t)y * bw + x);
      for (int k = 0; k < 16; k++) {
        const double v = (double)coef[k] * qstep * pre[k];
        out[k] = v >= 0.0 ? (int)(v + 0.5) : -(int)(-v + 0.5);  // ROUND (defines.h:23)
      }
    }
  *bitpos = b.pos;
  return DVC_OK;
}

}  // namespace dvc

extern "C" int dvc_intra_parse(const uint8_t *data, size_t len, size_t *bit_offset, const uint8_t *block_mode,
                               int width, int height, int intra_qp, int32_t *coeffs) {
  using namespace dvc;
  DVC_CHECK((data || len == 0) && bit_offset && block_mode && coeffs, DVC_EINVAL, "dvc_intra_parse: NULL argument");
  DVC_CHECK((width == 176 && height == 144) || (width == 352 && height == 288), DVC_EINVAL,
            "dvc_intra_parse: size must be QCIF or CIF");
  DVC_CHECK(intra_qp >= 0 && intra_qp <= 51, DVC_EINVAL, "dvc_intra_parse: intra_qp must be 0..51");
  return intra_parse_frame(data, len, bit_offset, block_mode, width, height, intra_qp, coeffs);
}
