// image_io.cpp -- output writers behind the C ABI (host code only), for the
// post-processed or linear RGB float images the renderer produces:
//   PPM16  OfflineEngine::coreLoop's output.ppm (src/core/OfflineEngine.cpp:80-137):
//          "P6\nW H\n65535\n", pixels from the last index to the first,
//          big-endian uint16 channels, value * 65535 truncated
//   PNG    saveImage for PNG (src/materials/Texture.h:44-60): 8-bit RGB, rows
//          from the top, clamp to [0, 1] then value * 255 truncated; written
//          with stored (uncompressed) deflate blocks
//   EXR    saveImage for EXR (Texture.h:62-69, tinyexr SaveEXR with FP32
//          channels): single-part scanline OpenEXR, channels B, G, R as FLOAT,
//          no compression
//   PFM    Portable FloatMap of linear radiance: "PF", rows bottom to top,
//          little-endian
#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ne_gpu.h"

namespace ne {
ne_status host_fail(ne_status st, const std::string& msg);
}

namespace {

class Out {
 public:
  void u8(uint8_t v) { b_.push_back(v); }
  void bytes(const void* p, size_t n) {
    const uint8_t* q = static_cast<const uint8_t*>(p);
    b_.insert(b_.end(), q, q + n);
  }
  void str(const std::string& s) { bytes(s.data(), s.size()); }
  void cstr(const char* s) { bytes(s, std::strlen(s) + 1); }
  void le32(uint32_t v) { for (int k = 0; k < 4; ++k) u8(uint8_t(v >> (8 * k))); }
  void le64(uint64_t v) { for (int k = 0; k < 8; ++k) u8(uint8_t(v >> (8 * k))); }
  void be16(uint16_t v) { u8(uint8_t(v >> 8)); u8(uint8_t(v)); }
  void be32(uint32_t v) { for (int k = 3; k >= 0; --k) u8(uint8_t(v >> (8 * k))); }
  void lef(float f) { uint32_t u; std::memcpy(&u, &f, 4); le32(u); }
  const std::vector<uint8_t>& data() const { return b_; }
  size_t size() const { return b_.size(); }

 private:
  std::vector<uint8_t> b_;
};

float clamp01(float v) { return v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f; }  // NaN -> 0

void ppm16(Out& o, int W, int H, const float* rgb) {
  o.str("P6\n" + std::to_string(W) + " " + std::to_string(H) + "\n65535\n");
  for (size_t i = size_t(W) * size_t(H); i-- > 0;)
    for (int c = 0; c < 3; ++c) o.be16(uint16_t(clamp01(rgb[3 * i + size_t(c)]) * 65535.0f));
}

uint32_t crc32(const uint8_t* p, size_t n, uint32_t crc = 0) {
  static const std::array<uint32_t, 256> table = [] {  // thread-safe one-time init
    std::array<uint32_t, 256> t{};
    for (uint32_t k = 0; k < 256; ++k) {
      uint32_t c = k;
      for (int j = 0; j < 8; ++j) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      t[k] = c;
    }
    return t;
  }();
  crc = ~crc;
  for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFF] ^ (crc >> 8);
  return ~crc;
}

void png_chunk(Out& o, const char* type, const std::vector<uint8_t>& body) {
  o.be32(uint32_t(body.size()));
  std::vector<uint8_t> tb(type, type + 4);
  tb.insert(tb.end(), body.begin(), body.end());
  o.bytes(tb.data(), tb.size());
  o.be32(crc32(tb.data(), tb.size()));
}

void png(Out& o, int W, int H, const float* rgb) {
  static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
  o.bytes(sig, 8);
  Out ih;
  ih.be32(uint32_t(W));
  ih.be32(uint32_t(H));
  ih.u8(8);  // bit depth
  ih.u8(2);  // colour type RGB
  ih.u8(0);  // deflate
  ih.u8(0);  // filter method
  ih.u8(0);  // no interlace
  png_chunk(o, "IHDR", ih.data());
  // raw scanlines: filter type 0 + RGB bytes
  std::vector<uint8_t> raw;
  raw.reserve(size_t(H) * (size_t(W) * 3 + 1));
  for (int y = 0; y < H; ++y) {
    raw.push_back(0);
    for (int x = 0; x < W; ++x)
      for (int c = 0; c < 3; ++c)
        raw.push_back(uint8_t(clamp01(rgb[(size_t(y) * size_t(W) + size_t(x)) * 3 + size_t(c)]) * 255.0f));
  }
  // zlib stream of stored deflate blocks (RFC 1950 / 1951)
  Out z;
  z.u8(0x78);
  z.u8(0x01);
  size_t pos = 0;
  do {
    const size_t n = std::min<size_t>(65535, raw.size() - pos);
    const bool last = pos + n == raw.size();
    z.u8(last ? 1 : 0);
    z.u8(uint8_t(n));
    z.u8(uint8_t(n >> 8));
    z.u8(uint8_t(~n));
    z.u8(uint8_t(~n >> 8));
    z.bytes(raw.data() + pos, n);
    pos += n;
  } while (pos < raw.size());
  uint32_t a = 1, b = 0;  // Adler-32
  for (uint8_t v : raw) {
    a = (a + v) % 65521u;
    b = (b + a) % 65521u;
  }
  z.be32((b << 16) | a);
  png_chunk(o, "IDAT", z.data());
  png_chunk(o, "IEND", {});
}

void exr_attr(Out& o, const char* name, const char* type, const Out& value) {
  o.cstr(name);
  o.cstr(type);
  o.le32(uint32_t(value.size()));
  o.bytes(value.data().data(), value.size());
}

void exr(Out& o, int W, int H, const float* rgb) {
  o.le32(20000630u);  // magic 76 2f 31 01
  o.le32(2u);         // version 2, single-part scanline
  Out ch;
  for (const char* n : {"B", "G", "R"}) {  // alphabetical channel order
    ch.cstr(n);
    ch.le32(2u);  // FLOAT
    ch.u8(0);     // pLinear
    ch.u8(0); ch.u8(0); ch.u8(0);
    ch.le32(1u);  // xSampling
    ch.le32(1u);  // ySampling
  }
  ch.u8(0);
  exr_attr(o, "channels", "chlist", ch);
  Out comp;
  comp.u8(0);  // NO_COMPRESSION
  exr_attr(o, "compression", "compression", comp);
  Out box;
  box.le32(0); box.le32(0); box.le32(uint32_t(W - 1)); box.le32(uint32_t(H - 1));
  exr_attr(o, "dataWindow", "box2i", box);
  exr_attr(o, "displayWindow", "box2i", box);
  Out lo;
  lo.u8(0);  // INCREASING_Y
  exr_attr(o, "lineOrder", "lineOrder", lo);
  Out par;
  par.lef(1.0f);
  exr_attr(o, "pixelAspectRatio", "float", par);
  Out swc;
  swc.lef(0.0f); swc.lef(0.0f);
  exr_attr(o, "screenWindowCenter", "v2f", swc);
  Out sww;
  sww.lef(1.0f);
  exr_attr(o, "screenWindowWidth", "float", sww);
  o.u8(0);  // end of header
  const uint64_t block = 8 + uint64_t(W) * 3 * 4;
  const uint64_t first = o.size() + uint64_t(H) * 8;
  for (int y = 0; y < H; ++y) o.le64(first + uint64_t(y) * block);
  for (int y = 0; y < H; ++y) {
    o.le32(uint32_t(y));
    o.le32(uint32_t(W * 3 * 4));
    for (int c : {2, 1, 0})
      for (int x = 0; x < W; ++x) o.lef(rgb[(size_t(y) * size_t(W) + size_t(x)) * 3 + size_t(c)]);
  }
}

void pfm(Out& o, int W, int H, const float* rgb) {
  o.str("PF\n" + std::to_string(W) + " " + std::to_string(H) + "\n-1.0\n");
  for (int y = H - 1; y >= 0; --y)
    for (int x = 0; x < W; ++x)
      for (int c = 0; c < 3; ++c) o.lef(rgb[(size_t(y) * size_t(W) + size_t(x)) * 3 + size_t(c)]);
}

}  // namespace

extern "C" ne_status ne_image_write(const char* path, int32_t format, int32_t width, int32_t height,
                                    const float* rgb) {
  if (!path || !rgb) return ne::host_fail(NE_ERR_INVALID_ARG, "path and pixels must not be NULL");
  if (width < 1 || height < 1 || width > 65535 || height > 65535)
    return ne::host_fail(NE_ERR_INVALID_ARG, "image size out of [1, 65535]");
  Out o;
  try {
    switch (format) {
      case NE_IMAGE_PPM16: ppm16(o, width, height, rgb); break;
      case NE_IMAGE_PNG: png(o, width, height, rgb); break;
      case NE_IMAGE_EXR: exr(o, width, height, rgb); break;
      case NE_IMAGE_PFM: pfm(o, width, height, rgb); break;
      default: return ne::host_fail(NE_ERR_INVALID_ARG, "unknown image format");
    }
  } catch (const std::bad_alloc&) {
    return ne::host_fail(NE_ERR_OUT_OF_MEMORY, "out of host memory while encoding the image");
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return ne::host_fail(NE_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
  const size_t n = std::fwrite(o.data().data(), 1, o.size(), f);
  const int rc = std::fclose(f);
  if (n != o.size() || rc != 0) return ne::host_fail(NE_ERR_IO, std::string("write error on ") + path);
  return NE_OK;
}
