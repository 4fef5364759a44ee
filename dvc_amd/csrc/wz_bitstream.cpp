// wz_bitstream.cpp -- the WZ bitstream reader: readWZBitstream
// (wzbitstream.cpp:128-210) and readGOP (wzbitstream.cpp:13-126) over a
// bitstream held in memory.  It yields, per WZ frame in the stream's order
// (which is decodeGOP's decoding order), the fields dvc_wz_frame takes: AC
// ranges, the block-mode maps decoded from their JBIG images (jbig_dec.cpp),
// the mode bits, and each bitplane's CRC and accumulated syndrome.
//
// Host code: parsing a file is sequential byte work done once per stream.
#include <cstring>
#include <vector>
#include "dvc_internal.h"

namespace dvc {
int jbig_decode_bie(const uint8_t *bie, size_t len, std::vector<uint8_t> &img, uint32_t &xd, uint32_t &yd,
                    size_t &consumed);
int intra_parse_frame(const uint8_t *data, size_t len, size_t *bitpos, const uint8_t *mode, int W, int H, int qp,
                      int32_t *coeffs);
}

struct dvc_wz_stream {
  dvc_wz_stream_info info{};
  struct Frame {
    int past, cur, fut;
    int acRange[16];
    int skipMode, intraMode, modeBits;
    std::vector<uint8_t> mode, crc, acc;
    std::vector<int32_t> intra;  // parsed intra coefficients (dvc_wz_stream_attach_intra)
  };
  bool intraAttached = false;
  std::vector<Frame> frames;
};

namespace dvc {
namespace {

// BlockModeFlag values (defines.h:15-17)
constexpr uint8_t kIntraB = 1, kSkipB = 2;

struct Reader {
  const uint8_t *d;
  size_t len, pos = 0;
  bool take(void *out, size_t n) {
    if (n > len - pos) return false;
    memcpy(out, d + pos, n);
    pos += n;
    return true;
  }
};

// One block-mode map (readGOP's JBIG branch, wzbitstream.cpp:40-66): the
// reference reads `bufferSize` bytes, decodes the map from them behind the
// header it rebuilds, and steps back to just after the bytes the decoder
// used.  Sets value `mark` in `mode` where the map has a 1 and returns the
// map's bit count (decoded bytes * 8 + 1, the mode flag included).
int read_mode_map(Reader &r, const dvc_wz_stream_info &in, uint8_t mark, std::vector<uint8_t> &mode, int *bits) {
  const int w8 = in.width / 8, h8 = in.height / 8;
  const size_t bufferSize = (size_t)((((w8 + 4) >> 3) << 3) * h8 >> 3) * 2;
  const size_t avail = std::min(bufferSize, r.len - r.pos);
  std::vector<uint8_t> bie(20 + avail);
  uint8_t *h = bie.data();
  // jbg_enc_init(w8, h8, 1 plane) + jbg_enc_options(ILEAVE|SMID, TPDON|DPON,
  // stripe h8, mx 8, my 0) + jbg_set_header
  memset(h, 0, 20);
  h[2] = 1;
  for (int k = 0; k < 4; k++) {
    h[4 + k] = (uint8_t)((uint32_t)w8 >> (24 - 8 * k));
    h[8 + k] = (uint8_t)((uint32_t)h8 >> (24 - 8 * k));
    h[12 + k] = (uint8_t)((uint32_t)h8 >> (24 - 8 * k));
  }
  h[16] = 8; h[18] = 0x03; h[19] = 0x14;
  memcpy(h + 20, r.d + r.pos, avail);
  std::vector<uint8_t> img;
  uint32_t xd = 0, yd = 0;
  size_t used = 0;
  int rc = jbig_decode_bie(bie.data(), bie.size(), img, xd, yd, used);
  if (rc) return rc;
  DVC_CHECK(xd == (uint32_t)w8 && yd == (uint32_t)h8, DVC_EINVAL, "WZ bitstream: block-mode map of the wrong size");
  const size_t bytes = used - 20;
  r.pos += bytes;
  for (size_t k = 0; k < img.size(); k++)
    if (img[k]) mode[k] = mark;
  *bits = (int)(bytes << 3) + 1;
  return DVC_OK;
}

int read_gop(Reader &r, dvc_wz_stream &s, int begin, int end) {
  const int cur = (begin + end) / 2;
  if (cur <= begin) return DVC_OK;
  const dvc_wz_stream_info &in = s.info;
  dvc_wz_stream::Frame f{};
  f.past = begin; f.cur = cur; f.fut = end;
  for (int j = 1; j < 16; j++) {  // AC ranges, 2 bytes each, in zig-zag order while Qtable > 0
    const int index = zigzag_entry(j);
    if (qtable_entry(in.qindex, index) <= 0) break;
    uint8_t b[2];
    DVC_CHECK(r.take(b, 2), DVC_EINVAL, "WZ bitstream: truncated AC range");
    f.acRange[index] = (b[0] << 8) | b[1];
  }
  const int n8 = (in.width / 8) * (in.height / 8);
  f.mode.assign(n8, 0);
  uint8_t flag = 0;
  int skipBits = 1, intraBits = 1;
  DVC_CHECK(r.take(&flag, 1), DVC_EINVAL, "WZ bitstream: truncated skip flag");
  f.skipMode = flag;
  if (flag == 1) {
    int rc = read_mode_map(r, in, kSkipB, f.mode, &skipBits);
    if (rc) return rc;
  }
  DVC_CHECK(r.take(&flag, 1), DVC_EINVAL, "WZ bitstream: truncated intra flag");
  f.intraMode = flag;
  if (flag == 1) {  // the intra map is read second, so it wins where both maps are set
    int rc = read_mode_map(r, in, kIntraB, f.mode, &intraBits);
    if (rc) return rc;
  }
  f.modeBits = skipBits + intraBits;
  const int nbp = in.nbitplanes, synd = (in.width / 4) * (in.height / 4) / 8;
  f.crc.resize(nbp);
  f.acc.resize((size_t)nbp * synd);
  for (int k = 0; k < nbp; k++) {
    DVC_CHECK(r.take(&f.crc[k], 1) && r.take(&f.acc[(size_t)k * synd], synd), DVC_EINVAL,
              "WZ bitstream: truncated bitplane");
  }
  s.frames.push_back(std::move(f));
  int rc = read_gop(r, s, begin, cur);
  if (rc) return rc;
  return read_gop(r, s, cur, end);
}

}  // namespace
}  // namespace dvc

using namespace dvc;

extern "C" int dvc_wz_stream_open(const uint8_t *data, size_t len, dvc_wz_stream **out) {
  DVC_CHECK(data && out, DVC_EINVAL, "dvc_wz_stream_open: NULL argument");
  *out = nullptr;
  Reader r{data, len};
  uint8_t h[7];
  DVC_CHECK(r.take(h, 7), DVC_EINVAL, "WZ bitstream: truncated header");
  auto *s = new dvc_wz_stream();
  dvc_wz_stream_info &in = s->info;
  if (h[0] == 0) { in.width = 176; in.height = 144; }
  else if (h[0] == 1) { in.width = 352; in.height = 288; }
  else { delete s; DVC_CHECK(false, DVC_EINVAL, "WZ bitstream: unsupported resolution (QCIF or CIF)"); }
  in.qindex = h[1];
  in.intra_qp = h[2];
  in.gop = h[3];
  in.frames = (h[4] << 8) | h[5];
  in.deblocking = h[6];
  if (in.qindex < 1 || in.qindex > 8 || in.gop < 1) {
    delete s;
    DVC_CHECK(false, DVC_EINVAL, "WZ bitstream: Qindex must be 1..8 and the GOP at least 1");
  }
  in.nbitplanes = 0;
  for (int j = 0; j < 16; j++) in.nbitplanes += qbits_entry(in.qindex, j);
  for (int i = 0; i < in.frames; i += in.gop) {
    const int end = i + in.gop >= in.frames ? in.frames - 1 : i + in.gop;
    const int rc = read_gop(r, *s, i, end);
    if (rc) { delete s; return rc; }
  }
  in.wz_frames = (int)s->frames.size();
  in.bytes_read = r.pos;
  *out = s;
  return DVC_OK;
}

extern "C" int dvc_wz_stream_info_get(const dvc_wz_stream *s, dvc_wz_stream_info *info) {
  DVC_CHECK(s && info, DVC_EINVAL, "dvc_wz_stream_info_get: NULL argument");
  *info = s->info;
  return DVC_OK;
}

extern "C" int dvc_wz_stream_attach_intra(dvc_wz_stream *s, const uint8_t *data, size_t len) {
  DVC_CHECK(s && (data || len == 0), DVC_EINVAL, "dvc_wz_stream_attach_intra: bad argument");
  DVC_CHECK(!s->intraAttached, DVC_ESTATE, "dvc_wz_stream_attach_intra: already attached");
  const dvc_wz_stream_info &in = s->info;
  size_t bit = 0;  // IntraBS.frame_bitoffset runs on from frame to frame
  for (auto &f : s->frames) {
    if (f.intraMode != 1) continue;
    f.intra.assign((size_t)(in.width / 4) * (in.height / 4) * 16, 0);
    const size_t before = bit;
    const int rc = intra_parse_frame(data, len, &bit, f.mode.data(), in.width, in.height, in.intra_qp, f.intra.data());
    if (rc) return rc;
    DVC_CHECK(bit <= len * 8, DVC_EINVAL, "WZ bitstream: the intra bitstream ends inside a frame's blocks");
    f.modeBits += (int)(bit - before);
  }
  s->intraAttached = true;
  return DVC_OK;
}

extern "C" int dvc_wz_stream_frame_get(const dvc_wz_stream *s, int k, dvc_wz_stream_frame *f) {
  DVC_CHECK(s && f && k >= 0 && k < (int)s->frames.size(), DVC_EINVAL, "dvc_wz_stream_frame_get: bad argument");
  const dvc_wz_stream::Frame &x = s->frames[k];
  bool intraBlocks = false;
  for (uint8_t m : x.mode) intraBlocks |= m == kIntraB;
  DVC_CHECK(!intraBlocks || s->intraAttached, DVC_ESTATE,
            "dvc_wz_stream_frame_get: the frame has intra blocks; attach the intra bitstream first");
  f->intra_coeffs = intraBlocks ? x.intra.data() : nullptr;
  // the past reference: a WZ frame of the stream when one has that index
  f->past_is_wz = 0;
  f->past_block_mode = nullptr;
  for (const auto &y : s->frames)
    if (y.cur == x.past) {
      f->past_is_wz = 1;
      f->past_block_mode = y.mode.data();
    }
  f->past_idx = x.past; f->cur_idx = x.cur; f->fut_idx = x.fut;
  memcpy(f->ac_range, x.acRange, sizeof(f->ac_range));
  f->skip_mode = x.skipMode;
  f->intra_mode = x.intraMode;
  f->mode_bits = x.modeBits;
  f->block_mode = x.mode.data();
  f->crc = x.crc.data();
  f->acc = x.acc.data();
  return DVC_OK;
}

extern "C" int dvc_wz_stream_close(dvc_wz_stream *s) {
  delete s;
  return DVC_OK;
}
