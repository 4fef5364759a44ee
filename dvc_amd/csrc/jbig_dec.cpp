// jbig_dec.cpp -- decoder of the JBIG (ITU-T T.82) images that carry the WZ
// bitstream's block-mode maps: the skip map and the intra map of a WZ frame,
// one bit per 8x8 block (wzbitstream.cpp:38-101 reads each with
// simple_JBIG_dec, sJBIG.cpp:49-103, over the vendored JBIG-KIT, and unpacks
// it with ByteToBitmap, sJBIG.cpp:24-47).
//
// Host code: a map is at most 44 x 36 pixels, decoded once per WZ frame by an
// adaptive arithmetic decoder whose every step depends on the previous one.
//
// Scope: what a bi-level image entity (BIE) of this kind can hold -- one bit
// plane and only the lowest-resolution layer (D = 0, so the resolution
// reduction, deterministic prediction and differential typical prediction of
// T.82 never apply; a private DP table, if signalled, is skipped), any stripe
// height L0, the two- and three-line templates, typical prediction of the
// lowest layer (TPBON), adaptive-template moves within the row (ATMOVE with
// ty = 0; MY must be 0), NEWLEN (with VLENGTH), COMMENT, and stripes ended by
// SDNORM (probability states carried to the next stripe) or SDRST (states and
// the lines above reset).  Anything else is reported as an error.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>
#include "dvc_internal.h"

namespace dvc {
namespace {

// The QM-coder's probability estimation state machine (T.82 Table 24, the
// same table as T.81 Table D.2): LPS sub-interval size, next state after an
// MPS renormalisation, next state after an LPS, and whether the LPS swaps
// the MPS sense.
struct QmEntry {
  uint16_t lsz;
  uint8_t nmps, nlps, swap;
};
constexpr QmEntry kQm[113] = {
    {0x5a1d, 1, 1, 1},     {0x2586, 2, 14, 0},    {0x1114, 3, 16, 0},    {0x080b, 4, 18, 0},
    {0x03d8, 5, 20, 0},    {0x01da, 6, 23, 0},    {0x00e5, 7, 25, 0},    {0x006f, 8, 28, 0},
    {0x0036, 9, 30, 0},    {0x001a, 10, 33, 0},   {0x000d, 11, 35, 0},   {0x0006, 12, 9, 0},
    {0x0003, 13, 10, 0},   {0x0001, 13, 12, 0},   {0x5a7f, 15, 15, 1},   {0x3f25, 16, 36, 0},
    {0x2cf2, 17, 38, 0},   {0x207c, 18, 39, 0},   {0x17b9, 19, 40, 0},   {0x1182, 20, 42, 0},
    {0x0cef, 21, 43, 0},   {0x09a1, 22, 45, 0},   {0x072f, 23, 46, 0},   {0x055c, 24, 48, 0},
    {0x0406, 25, 49, 0},   {0x0303, 26, 51, 0},   {0x0240, 27, 52, 0},   {0x01b1, 28, 54, 0},
    {0x0144, 29, 56, 0},   {0x00f5, 30, 57, 0},   {0x00b7, 31, 59, 0},   {0x008a, 32, 60, 0},
    {0x0068, 33, 62, 0},   {0x004e, 34, 63, 0},   {0x003b, 35, 32, 0},   {0x002c, 9, 33, 0},
    {0x5ae1, 37, 37, 1},   {0x484c, 38, 64, 0},   {0x3a0d, 39, 65, 0},   {0x2ef1, 40, 67, 0},
    {0x261f, 41, 68, 0},   {0x1f33, 42, 69, 0},   {0x19a8, 43, 70, 0},   {0x1518, 44, 72, 0},
    {0x1177, 45, 73, 0},   {0x0e74, 46, 74, 0},   {0x0bfb, 47, 75, 0},   {0x09f8, 48, 77, 0},
    {0x0861, 49, 78, 0},   {0x0706, 50, 79, 0},   {0x05cd, 51, 48, 0},   {0x04de, 52, 50, 0},
    {0x040f, 53, 50, 0},   {0x0363, 54, 51, 0},   {0x02d4, 55, 52, 0},   {0x025c, 56, 53, 0},
    {0x01f8, 57, 54, 0},   {0x01a4, 58, 55, 0},   {0x0160, 59, 56, 0},   {0x0125, 60, 57, 0},
    {0x00f6, 61, 58, 0},   {0x00cb, 62, 59, 0},   {0x00ab, 63, 61, 0},   {0x008f, 32, 61, 0},
    {0x5b12, 65, 65, 1},   {0x4d04, 66, 80, 0},   {0x412c, 67, 81, 0},   {0x37d8, 68, 82, 0},
    {0x2fe8, 69, 83, 0},   {0x293c, 70, 84, 0},   {0x2379, 71, 86, 0},   {0x1edf, 72, 87, 0},
    {0x1aa9, 73, 87, 0},   {0x174e, 74, 72, 0},   {0x1424, 75, 72, 0},   {0x119c, 76, 74, 0},
    {0x0f6b, 77, 74, 0},   {0x0d51, 78, 75, 0},   {0x0bb6, 79, 77, 0},   {0x0a40, 48, 77, 0},
    {0x5832, 81, 80, 1},   {0x4d1c, 82, 88, 0},   {0x438e, 83, 89, 0},   {0x3bdd, 84, 90, 0},
    {0x34ee, 85, 91, 0},   {0x2eae, 86, 92, 0},   {0x299a, 87, 93, 0},   {0x2516, 71, 86, 0},
    {0x5570, 89, 88, 1},   {0x4ca9, 90, 95, 0},   {0x44d9, 91, 96, 0},   {0x3e22, 92, 97, 0},
    {0x3824, 93, 99, 0},   {0x32b4, 94, 99, 0},   {0x2e17, 86, 93, 0},   {0x56a8, 96, 95, 1},
    {0x4f46, 97, 101, 0},  {0x47e5, 98, 102, 0},  {0x41cf, 99, 103, 0},  {0x3c3d, 100, 104, 0},
    {0x375e, 93, 99, 0},   {0x5231, 102, 105, 0}, {0x4c0f, 103, 106, 0}, {0x4639, 104, 107, 0},
    {0x415e, 99, 103, 0},  {0x5627, 106, 105, 1}, {0x50e7, 107, 108, 0}, {0x4b85, 103, 109, 0},
    {0x5597, 109, 110, 0}, {0x504f, 107, 111, 0}, {0x5a10, 111, 110, 1}, {0x5522, 109, 112, 0},
    {0x59eb, 111, 112, 1},
};

// Adaptive binary arithmetic decoder over one stripe's coded data (stuffing
// already removed).  Past the end of the data it reads zero bytes, as T.82
// has the decoder do after the stripe's end marker.  C keeps the code value
// in its upper 16 bits and up to 16 further bits below; a byte is shifted in
// whenever the lower half runs dry.
class QmDecoder {
 public:
  void start(const uint8_t *b, const uint8_t *e) {
    p_ = b;
    end_ = e;
    c_ = next() << 24;
    c_ |= next() << 16;
    a_ = 0x10000;
    ct_ = 0;
  }
  // One decision in the context whose state byte is st (bit 7: the MPS,
  // bits 0-6: the estimation state).
  int decode(uint8_t &st) {
    const int s = st & 0x7f, mps = st >> 7;
    const uint32_t lsz = kQm[s].lsz;
    a_ -= lsz;
    int d;
    bool lpsTransition;
    if ((c_ >> 16) < a_) {
      if (a_ >= 0x8000) return mps;  // MPS, interval still normalised
      lpsTransition = a_ < lsz;      // conditional exchange
      d = lpsTransition ? 1 - mps : mps;
    } else {
      c_ -= a_ << 16;
      lpsTransition = a_ >= lsz;
      d = lpsTransition ? 1 - mps : mps;
      a_ = lsz;
    }
    if (lpsTransition)
      st = (uint8_t)(((kQm[s].swap ? 1 - mps : mps) << 7) | kQm[s].nlps);
    else
      st = (uint8_t)((mps << 7) | kQm[s].nmps);
    do {
      if (ct_ == 0) {
        c_ |= next() << 8;
        ct_ = 8;
      }
      a_ <<= 1;
      c_ <<= 1;
      ct_--;
    } while (a_ < 0x8000);
    return d;
  }

 private:
  uint32_t next() { return p_ < end_ ? *p_++ : 0u; }
  const uint8_t *p_ = nullptr, *end_ = nullptr;
  uint32_t c_ = 0, a_ = 0;
  int ct_ = 0;
};

constexpr uint8_t kEsc = 0xff, kStuff = 0x00, kSdnorm = 0x02, kSdrst = 0x03, kAbort = 0x04, kNewlen = 0x05,
                  kAtmove = 0x06, kComment = 0x07;
constexpr uint8_t kOptLrltwo = 0x40, kOptVlength = 0x20, kOptTpbon = 0x08, kOptDpon = 0x04, kOptDppriv = 0x02,
                  kOptDplast = 0x01;
constexpr int kTpbCx3 = 0x0e5, kTpbCx2 = 0x195;  // contexts of the typical-prediction flag

uint32_t be32(const uint8_t *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}

}  // namespace

// Decodes one BIE (20-byte header + data) of a single-plane, single-layer
// image into `img` (one byte 0/1 per pixel, row-major, xd * yd).  *consumed:
// bytes of the BIE read through the end marker of its last stripe.
int jbig_decode_bie(const uint8_t *bie, size_t len, std::vector<uint8_t> &img, uint32_t &xd, uint32_t &yd,
                    size_t &consumed) {
  DVC_CHECK(bie && len >= 20, DVC_EINVAL, "jbig: the header needs 20 bytes");
  const uint8_t *h = bie;
  const int dl = h[0], d = h[1], planes = h[2];
  xd = be32(h + 4);
  yd = be32(h + 8);
  const uint32_t l0 = be32(h + 12);
  const int mx = h[16], my = h[17], options = h[19];
  DVC_CHECK(h[3] == 0 && (h[18] & 0xf0) == 0 && (h[19] & 0x80) == 0, DVC_EINVAL, "jbig: header padding not zero");
  DVC_CHECK(dl == 0 && d == 0 && planes == 1, DVC_EINVAL,
            "jbig: only single-plane images of the lowest resolution layer are supported");
  DVC_CHECK(my == 0, DVC_EINVAL, "jbig: vertical adaptive-template offsets are not supported");
  DVC_CHECK(xd > 0 && yd > 0 && l0 > 0 && xd <= 65535 && yd <= 65535, DVC_EINVAL, "jbig: bad image size");
  size_t pos = 20;
  if ((options & (kOptDpon | kOptDppriv | kOptDplast)) == (kOptDpon | kOptDppriv)) {
    DVC_CHECK(len >= pos + 1728, DVC_EINVAL, "jbig: truncated private DP table");
    pos += 1728;  // only differential layers use it
  }
  const bool two = options & kOptLrltwo, tpbon = options & kOptTpbon;
  img.assign((size_t)xd * yd, 0);
  std::vector<uint8_t> st(4096, 0);  // context states
  std::vector<uint8_t> pscd;
  QmDecoder qm;
  int tx = 0;        // adaptive-template pixel: (x - tx, y), 0 = default position
  bool lntp = true;  // previous line was not typical
  bool reset = false;
  uint32_t y0 = 0;   // first line of the current stripe
  struct AtMove {
    uint32_t line;
    int tx;
  };
  std::vector<AtMove> moves;
  auto pix = [&](long x, long y, uint32_t top) -> int {
    if (x < 0 || x >= (long)xd || y < (long)top) return 0;
    return img[(size_t)y * xd + x];
  };
  // A floating marker segment at pos (ATMOVE, NEWLEN, COMMENT): applied and
  // skipped; returns 1 when one was there, 0 when pos holds something else,
  // or an error code.
  auto floating = [&](size_t &at) -> int {
    const uint8_t m = bie[at + 1];
    if (m == kAtmove) {
      DVC_CHECK(at + 8 <= len, DVC_EINVAL, "jbig: truncated ATMOVE");
      const int t = (int8_t)bie[at + 6];
      DVC_CHECK(bie[at + 7] == 0 && t >= 0 && t <= mx, DVC_EINVAL, "jbig: ATMOVE outside the allowed range");
      moves.push_back({be32(bie + at + 2), t});
      at += 8;
      return 1;
    }
    if (m == kNewlen) {
      DVC_CHECK(at + 6 <= len && (options & kOptVlength), DVC_EINVAL, "jbig: unexpected NEWLEN");
      const uint32_t ny = be32(bie + at + 2);
      DVC_CHECK(ny <= yd && ny >= y0, DVC_EINVAL, "jbig: NEWLEN beyond the image");
      yd = ny;
      at += 6;
      return 1;
    }
    if (m == kComment) {
      DVC_CHECK(at + 6 <= len, DVC_EINVAL, "jbig: truncated COMMENT");
      const size_t n = be32(bie + at + 2);
      DVC_CHECK(n <= len - at - 6, DVC_EINVAL, "jbig: truncated COMMENT");
      at += 6 + n;
      return 1;
    }
    DVC_CHECK(m != kAbort, DVC_EINVAL, "jbig: ABORT marker");
    return 0;
  };
  while (true) {
    // floating marker segments ahead of the stripe's coded data
    while (pos + 1 < len && bie[pos] == kEsc && bie[pos + 1] != kStuff) {
      const int r = floating(pos);
      if (r < 0) return r;
      if (r == 0) break;  // the stripe's end marker: a stripe with no coded bytes
    }
    if (y0 >= yd) break;  // a NEWLEN ended the image
    // the stripe's coded data: up to its end marker, with 0xff 0x00 -> 0xff
    // (floating marker segments inside it are applied and skipped)
    pscd.clear();
    while (true) {
      DVC_CHECK(pos < len, DVC_EINVAL, "jbig: data ends inside a stripe");
      if (bie[pos] != kEsc) {
        pscd.push_back(bie[pos++]);
        continue;
      }
      DVC_CHECK(pos + 1 < len, DVC_EINVAL, "jbig: data ends inside a marker");
      if (bie[pos + 1] == kStuff) {
        pscd.push_back(kEsc);
        pos += 2;
        continue;
      }
      const int r = floating(pos);
      if (r < 0) return r;
      if (r == 0) break;
    }
    const uint8_t endMarker = bie[pos + 1];
    DVC_CHECK(endMarker == kSdnorm || endMarker == kSdrst, DVC_EINVAL,
              "jbig: stripe data not ended by SDNORM or SDRST");
    pos += 2;
    // decode the stripe's lines
    const uint32_t y1 = (uint32_t)std::min<uint64_t>((uint64_t)y0 + l0, yd);
    if (y0 == 0 || reset) {
      tx = 0;
      lntp = true;
    }
    const uint32_t top = reset ? y0 : 0;  // lines above this one read as 0
    qm.start(pscd.data(), pscd.data() + pscd.size());
    for (uint32_t y = y0; y < y1; y++) {
      for (const AtMove &m : moves)
        if (m.line == y - y0) tx = m.tx;
      if (tpbon) {
        lntp = !(qm.decode(st[two ? kTpbCx2 : kTpbCx3]) ^ lntp);
        if (!lntp) {  // a typical line repeats the one above
          for (uint32_t x = 0; x < xd; x++) img[(size_t)y * xd + x] = (uint8_t)pix(x, (long)y - 1, top);
          continue;
        }
      }
      for (long x = 0; x < (long)xd; x++) {
        const long yy = (long)y;
        const int at = tx ? pix(x - tx, yy, top) : pix(x + 2, yy - 1, top);
        int cx;
        if (two)
          cx = (pix(x - 3, yy - 1, top) << 9) | (pix(x - 2, yy - 1, top) << 8) | (pix(x - 1, yy - 1, top) << 7) |
               (pix(x, yy - 1, top) << 6) | (pix(x + 1, yy - 1, top) << 5) | (at << 4) |
               (pix(x - 4, yy, top) << 3) | (pix(x - 3, yy, top) << 2) | (pix(x - 2, yy, top) << 1) |
               pix(x - 1, yy, top);
        else
          cx = (pix(x - 1, yy - 2, top) << 9) | (pix(x, yy - 2, top) << 8) | (pix(x + 1, yy - 2, top) << 7) |
               (pix(x - 2, yy - 1, top) << 6) | (pix(x - 1, yy - 1, top) << 5) | (pix(x, yy - 1, top) << 4) |
               (pix(x + 1, yy - 1, top) << 3) | (at << 2) | (pix(x - 2, yy, top) << 1) | pix(x - 1, yy, top);
        img[(size_t)y * xd + x] = (uint8_t)qm.decode(st[cx]);
      }
    }
    moves.clear();
    reset = endMarker == kSdrst;
    if (reset) std::fill(st.begin(), st.end(), 0);
    y0 = y1;
    if (y0 >= yd) break;
  }
  consumed = pos;
  return DVC_OK;
}

}  // namespace dvc

using namespace dvc;

extern "C" int dvc_jbig_decode_map(const uint8_t *data, size_t len, int width, int height, uint8_t *map,
                                   size_t *consumed) {
  DVC_CHECK(data && map && width > 0 && height > 0 && width <= 65535 && height <= 65535, DVC_EINVAL,
            "dvc_jbig_decode_map: bad argument");
  // The header the reference rebuilds for a map (wzbitstream.cpp:43-52):
  // jbg_enc_init(width, height, 1 plane), then order ILEAVE|SMID, options
  // TPDON|DPON, one stripe of `height` lines, MX 8, MY 0.
  std::vector<uint8_t> bie(20 + len);
  uint8_t *h = bie.data();
  h[0] = 0; h[1] = 0; h[2] = 1; h[3] = 0;
  for (int k = 0; k < 4; k++) {
    h[4 + k] = (uint8_t)((uint32_t)width >> (24 - 8 * k));
    h[8 + k] = (uint8_t)((uint32_t)height >> (24 - 8 * k));
    h[12 + k] = (uint8_t)((uint32_t)height >> (24 - 8 * k));
  }
  h[16] = 8; h[17] = 0; h[18] = 0x03; h[19] = 0x14;
  if (len) memcpy(h + 20, data, len);
  std::vector<uint8_t> img;
  uint32_t xd = 0, yd = 0;
  size_t used = 0;
  int rc = jbig_decode_bie(bie.data(), bie.size(), img, xd, yd, used);
  if (rc) return rc;
  DVC_CHECK(xd == (uint32_t)width && yd == (uint32_t)height, DVC_EINVAL, "dvc_jbig_decode_map: size changed");
  memcpy(map, img.data(), img.size());
  if (consumed) *consumed = used - 20;
  return DVC_OK;
}

extern "C" int dvc_jbig_decode(const uint8_t *bie, size_t len, uint8_t *image, size_t capacity, uint32_t *width,
                               uint32_t *height, size_t *consumed) {
  DVC_CHECK(bie && width && height, DVC_EINVAL, "dvc_jbig_decode: bad argument");
  std::vector<uint8_t> img;
  uint32_t xd = 0, yd = 0;
  size_t used = 0;
  int rc = jbig_decode_bie(bie, len, img, xd, yd, used);
  if (rc) return rc;
  *width = xd;
  *height = yd;
  if (consumed) *consumed = used;
  DVC_CHECK(image == nullptr || capacity >= img.size(), DVC_EINVAL, "dvc_jbig_decode: image buffer too small");
  if (image) memcpy(image, img.data(), img.size());
  return DVC_OK;
}
