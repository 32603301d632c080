// scene_io.cpp -- scene files behind the C ABI (host code only):
//   * a reader for the reference's JSON scene format, following SceneReader
//     (src/io/SceneReader.cpp:10-675): materials, primitives with their
//     instance transforms, camera and renderer settings, flattened into the
//     world-space ne_scene_desc that ne_scene_upload takes;
//   * the text .vol grid reader (ResourceManager::loadVolasTexture,
//     src/core/ResourceManager.cpp:222-286) and a raw little-endian .f32 grid
//     reader for large grids;
//   * a Wavefront OBJ reader for positions and faces (the reference imports
//     models through Assimp, src/primitives/Model.cpp:49-91; the path uses
//     only the triangles);
//   * Radiance .hdr and .pfm readers for InfiniteAreaLight maps (the
//     reference decodes images with stb_image).
// Failures are thrown internally as Failure and turned into an ne_status at
// the ABI; nothing throws across it.
#include <algorithm>
#include <cctype>
#include <charconv>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ne_gpu.h"

namespace ne {
ne_status host_fail(ne_status st, const std::string& msg);
}

namespace {

struct Failure {
  ne_status code;
  std::string msg;
};
[[noreturn]] void raise(ne_status code, const std::string& msg) { throw Failure{code, msg}; }

// ---------------------------------------------------------------------------
// JSON (RFC 8259, strict like rapidjson's defaults: no comments, no trailing
// commas).  Objects keep their member order; lookups return the first match.
// ---------------------------------------------------------------------------
struct Json {
  enum Kind { kNull, kBool, kNum, kStr, kArr, kObj } kind = kNull;
  bool b = false;
  double num = 0;
  std::string str;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;
  const Json* find(const char* key) const {
    if (kind != kObj) return nullptr;
    for (const auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

class JsonParser {
 public:
  JsonParser(const char* text, size_t n) : p_(text), e_(text + n) {}
  Json parse() {
    Json v;
    ws();
    value(v, 0);
    ws();
    if (p_ != e_) error("unexpected characters after the top-level value");
    return v;
  }

 private:
  const char* p_;
  const char* e_;
  int line_ = 1;

  [[noreturn]] void error(const std::string& what) {
    raise(NE_ERR_PARSE, "malformed JSON (line " + std::to_string(line_) + "): " + what);
  }
  void ws() {
    while (p_ < e_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\r' || *p_ == '\n')) {
      if (*p_ == '\n') ++line_;
      ++p_;
    }
  }
  void expect(char c) {
    if (p_ >= e_ || *p_ != c) error(std::string("expected '") + c + "'");
    ++p_;
  }
  void literal(const char* word) {
    size_t n = std::strlen(word);
    if (size_t(e_ - p_) < n || std::strncmp(p_, word, n) != 0) error("invalid literal");
    p_ += n;
  }
  void value(Json& v, int depth) {
    if (depth > 512) error("nesting too deep");
    if (p_ >= e_) error("unexpected end of input");
    const char c = *p_;
    if (c == '{') {
      v.kind = Json::kObj;
      ++p_;
      ws();
      if (p_ < e_ && *p_ == '}') { ++p_; return; }
      while (true) {
        ws();
        if (p_ >= e_ || *p_ != '"') error("expected a member name");
        std::string key;
        string(key);
        ws();
        expect(':');
        ws();
        v.obj.emplace_back(std::move(key), Json());
        value(v.obj.back().second, depth + 1);
        ws();
        if (p_ < e_ && *p_ == ',') { ++p_; continue; }
        expect('}');
        return;
      }
    }
    if (c == '[') {
      v.kind = Json::kArr;
      ++p_;
      ws();
      if (p_ < e_ && *p_ == ']') { ++p_; return; }
      while (true) {
        ws();
        v.arr.emplace_back();
        value(v.arr.back(), depth + 1);
        ws();
        if (p_ < e_ && *p_ == ',') { ++p_; continue; }
        expect(']');
        return;
      }
    }
    if (c == '"') { v.kind = Json::kStr; string(v.str); return; }
    if (c == 't') { literal("true"); v.kind = Json::kBool; v.b = true; return; }
    if (c == 'f') { literal("false"); v.kind = Json::kBool; v.b = false; return; }
    if (c == 'n') { literal("null"); v.kind = Json::kNull; return; }
    number(v);
  }
  static void utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) out += char(cp);
    else if (cp < 0x800) { out += char(0xC0 | (cp >> 6)); out += char(0x80 | (cp & 0x3F)); }
    else if (cp < 0x10000) {
      out += char(0xE0 | (cp >> 12)); out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F));
    } else {
      out += char(0xF0 | (cp >> 18)); out += char(0x80 | ((cp >> 12) & 0x3F));
      out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F));
    }
  }
  uint32_t hex4() {
    if (e_ - p_ < 4) error("truncated \\u escape");
    uint32_t v = 0;
    for (int i = 0; i < 4; ++i) {
      char h = *p_++;
      v <<= 4;
      if (h >= '0' && h <= '9') v |= uint32_t(h - '0');
      else if (h >= 'a' && h <= 'f') v |= uint32_t(h - 'a' + 10);
      else if (h >= 'A' && h <= 'F') v |= uint32_t(h - 'A' + 10);
      else error("invalid \\u escape");
    }
    return v;
  }
  void string(std::string& out) {
    expect('"');
    while (true) {
      if (p_ >= e_) error("unterminated string");
      const unsigned char c = static_cast<unsigned char>(*p_++);
      if (c == '"') return;
      if (c < 0x20) error("control character in a string");
      if (c != '\\') { out += char(c); continue; }
      if (p_ >= e_) error("unterminated escape");
      const char x = *p_++;
      switch (x) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          uint32_t cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00) {
            if (e_ - p_ < 6 || p_[0] != '\\' || p_[1] != 'u') error("unpaired surrogate");
            p_ += 2;
            uint32_t lo = hex4();
            if (lo < 0xDC00 || lo >= 0xE000) error("unpaired surrogate");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          } else if (cp >= 0xDC00 && cp < 0xE000) {
            error("unpaired surrogate");
          }
          utf8(out, cp);
          break;
        }
        default: error("invalid escape");
      }
    }
  }
  void number(Json& v) {
    const char* s = p_;
    auto digit = [&](const char* q) { return q < e_ && *q >= '0' && *q <= '9'; };
    if (p_ < e_ && *p_ == '-') ++p_;
    if (!digit(p_)) error("invalid value");
    if (*p_ == '0') ++p_;
    else while (digit(p_)) ++p_;
    if (p_ < e_ && *p_ == '.') {
      ++p_;
      if (!digit(p_)) error("invalid number");
      while (digit(p_)) ++p_;
    }
    if (p_ < e_ && (*p_ == 'e' || *p_ == 'E')) {
      ++p_;
      if (p_ < e_ && (*p_ == '+' || *p_ == '-')) ++p_;
      if (!digit(p_)) error("invalid number");
      while (digit(p_)) ++p_;
    }
    v.kind = Json::kNum;
    const auto r = std::from_chars(s, p_, v.num);  // locale-independent, like rapidjson
    if (r.ec == std::errc::result_out_of_range) error("number out of range");
  }
};

// ---------------------------------------------------------------------------
// Files
// ---------------------------------------------------------------------------
std::string read_file(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) raise(NE_ERR_IO, "Couldn't read " + path + ": " + std::strerror(errno));
  std::string out;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  if (bad) raise(NE_ERR_IO, "read error on " + path);
  return out;
}

std::string join_path(const std::string& base, const std::string& rel) {
  if (rel.empty() || rel[0] == '/' || base.empty()) return rel;
  return base.back() == '/' ? base + rel : base + "/" + rel;
}

std::string dir_of(const std::string& path) {
  size_t k = path.find_last_of('/');
  return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

bool ends_with(const std::string& s, const char* suffix) {
  const size_t n = std::strlen(suffix);
  return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

// Whitespace tokenizer over a text buffer that tracks line numbers.
struct Tokens {
  const char* p;
  const char* e;
  bool next(const char*& b, const char*& q) {
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == '\f' || *p == '\v')) ++p;
    if (p >= e) return false;
    b = p;
    while (p < e && !(*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == '\f' || *p == '\v')) ++p;
    q = p;
    return true;
  }
};

// Locale-independent float token (decimal point '.'), read as a double and
// rounded to float (values beyond the float range become +-inf or 0).
bool parse_float(const char* b, const char* q, float& out) {
  if (b < q && *b == '+') ++b;  // from_chars takes no leading '+'
  double d = 0.0;
  const auto r = std::from_chars(b, q, d);
  if (b >= q || r.ptr != q || (r.ec != std::errc() && r.ec != std::errc::result_out_of_range)) return false;
  out = float(d);
  return true;
}

// .vol text grid: line 1 holds the resolution "W H D", line 2 is skipped (as
// the reference's reader does), then W*H*D floats, x fastest.  The reference
// splits on single spaces and drops a last value that no space follows; any
// whitespace separates values here and every value is kept.
void read_vol(const std::string& path, int32_t res[3], std::vector<float>* grid) {
  const std::string text = read_file(path);
  const char* b = text.data();
  const char* e = b + text.size();
  const char* l1 = static_cast<const char*>(std::memchr(b, '\n', size_t(e - b)));
  const char* h_end = l1 ? l1 : e;
  Tokens th{b, h_end};
  const char *tb, *tq;
  double r[3];
  for (int a = 0; a < 3; ++a) {
    float v;
    if (!th.next(tb, tq) || !parse_float(tb, tq, v)) raise(NE_ERR_PARSE, path + ": the first line must hold W H D");
    r[a] = v;
  }
  for (int a = 0; a < 3; ++a) {
    if (!(r[a] >= 1 && r[a] <= 4096) || r[a] != std::floor(r[a]))
      raise(NE_ERR_PARSE, path + ": grid resolution must be integers in [1, 4096]");
    res[a] = int32_t(r[a]);
  }
  if (!grid) return;
  const char* l2 = l1 ? static_cast<const char*>(std::memchr(l1 + 1, '\n', size_t(e - l1 - 1))) : nullptr;
  const size_t n = size_t(res[0]) * size_t(res[1]) * size_t(res[2]);
  grid->assign(n, 0.0f);
  Tokens td{l2 ? l2 + 1 : e, e};
  size_t k = 0;
  while (td.next(tb, tq)) {
    if (k >= n) raise(NE_ERR_PARSE, path + ": more than W*H*D values");
    if (!parse_float(tb, tq, (*grid)[k])) raise(NE_ERR_PARSE, path + ": invalid value '" + std::string(tb, tq) + "'");
    ++k;
  }
  if (k != n)
    raise(NE_ERR_PARSE, path + ": " + std::to_string(k) + " values for a " + std::to_string(res[0]) + "x" +
                            std::to_string(res[1]) + "x" + std::to_string(res[2]) + " grid");
}

void read_f32(const std::string& path, const int32_t res[3], std::vector<float>& grid) {
  const std::string bytes = read_file(path);
  const size_t n = size_t(res[0]) * size_t(res[1]) * size_t(res[2]);
  if (bytes.size() != n * sizeof(float))
    raise(NE_ERR_PARSE, path + ": " + std::to_string(bytes.size()) + " bytes for a grid of " + std::to_string(n) +
                            " floats");
  grid.resize(n);
  std::memcpy(grid.data(), bytes.data(), bytes.size());
}

// ---------------------------------------------------------------------------
// Environment images (linear RGB float, row 0 at the top): Portable FloatMap
// (.pfm, "PF"/"Pf", rows stored bottom to top, negative scale = little
// endian) and Radiance RGBE (.hdr, flat or new-style run-length scanlines,
// value = mantissa * 2^(exponent - 136) as stb_image decodes it).
// ---------------------------------------------------------------------------
struct Image {
  int w = 0, h = 0;
  std::vector<float> rgb;
};

void read_pfm(const std::string& path, Image& img) {
  const std::string b = read_file(path);
  size_t p = 0;
  auto token = [&]() {
    while (p < b.size() && std::isspace(static_cast<unsigned char>(b[p]))) ++p;
    size_t s = p;
    while (p < b.size() && !std::isspace(static_cast<unsigned char>(b[p]))) ++p;
    return b.substr(s, p - s);
  };
  const std::string magic = token();
  if (magic != "PF" && magic != "Pf") raise(NE_ERR_PARSE, path + ": not a PFM file");
  const int ch = magic == "PF" ? 3 : 1;
  char* end = nullptr;
  const std::string ws = token(), hs = token(), ss = token();
  const long w = std::strtol(ws.c_str(), &end, 10);
  const long h = std::strtol(hs.c_str(), &end, 10);
  float scale = 0.0f;
  if (!parse_float(ss.data(), ss.data() + ss.size(), scale)) raise(NE_ERR_PARSE, path + ": bad PFM scale");
  if (w < 1 || h < 1 || w > 16384 || h > 16384 || scale == 0.0f) raise(NE_ERR_PARSE, path + ": bad PFM header");
  ++p;  // the single whitespace character after the scale
  const size_t n = size_t(w) * size_t(h) * size_t(ch);
  if (b.size() < p + 4 * n) raise(NE_ERR_PARSE, path + ": truncated PFM data");
  const bool little = scale < 0;
  img.w = int(w);
  img.h = int(h);
  img.rgb.assign(size_t(w) * size_t(h) * 3, 0.0f);
  for (long y = 0; y < h; ++y)
    for (long x = 0; x < w; ++x)
      for (int c = 0; c < 3; ++c) {
        const size_t src = p + 4 * ((size_t(y) * size_t(w) + size_t(x)) * size_t(ch) + size_t(ch == 3 ? c : 0));
        unsigned char q[4];
        std::memcpy(q, b.data() + src, 4);
        if (!little) { std::swap(q[0], q[3]); std::swap(q[1], q[2]); }
        float v;
        std::memcpy(&v, q, 4);
        img.rgb[((size_t(h - 1 - y)) * size_t(w) + size_t(x)) * 3 + size_t(c)] = v;  // bottom-up -> top-down
      }
}

void read_hdr(const std::string& path, Image& img) {
  const std::string b = read_file(path);
  size_t p = 0;
  auto line = [&]() {
    size_t e = b.find('\n', p);
    if (e == std::string::npos) raise(NE_ERR_PARSE, path + ": truncated Radiance header");
    std::string l = b.substr(p, e - p);
    p = e + 1;
    return l;
  };
  const std::string first = line();
  if (first.rfind("#?RADIANCE", 0) != 0 && first.rfind("#?RGBE", 0) != 0)
    raise(NE_ERR_PARSE, path + ": not a Radiance HDR file");
  while (true) {
    const std::string l = line();
    if (l.empty()) break;
    if (l.rfind("FORMAT=", 0) == 0 && l != "FORMAT=32-bit_rle_rgbe")
      raise(NE_ERR_UNSUPPORTED, path + ": only the 32-bit_rle_rgbe format is supported");
  }
  const std::string res = line();
  int h = 0, w = 0;
  char sy[3] = {0}, sx[3] = {0};
  if (std::sscanf(res.c_str(), "%2s %d %2s %d", sy, &h, sx, &w) != 4 || std::string(sy) != "-Y" ||
      std::string(sx) != "+X")
    raise(NE_ERR_UNSUPPORTED, path + ": only the standard -Y H +X W orientation is supported");
  if (w < 1 || h < 1 || w > 16384 || h > 16384) raise(NE_ERR_PARSE, path + ": bad image size");
  img.w = w;
  img.h = h;
  img.rgb.assign(size_t(w) * size_t(h) * 3, 0.0f);
  std::vector<unsigned char> scan(size_t(w) * 4);
  auto byte = [&]() -> unsigned char {
    if (p >= b.size()) raise(NE_ERR_PARSE, path + ": truncated pixel data");
    return static_cast<unsigned char>(b[p++]);
  };
  for (int y = 0; y < h; ++y) {
    const bool rle = w >= 8 && w < 32768 && p + 4 <= b.size() && static_cast<unsigned char>(b[p]) == 2 &&
                     static_cast<unsigned char>(b[p + 1]) == 2 && ((static_cast<unsigned char>(b[p + 2]) << 8) |
                                                                   static_cast<unsigned char>(b[p + 3])) == w;
    if (rle) {
      p += 4;
      for (int c = 0; c < 4; ++c) {
        int x = 0;
        while (x < w) {
          int count = byte();
          if (count > 128) {
            count -= 128;
            if (x + count > w) raise(NE_ERR_PARSE, path + ": run past the end of a scanline");
            const unsigned char v = byte();
            for (int k = 0; k < count; ++k) scan[size_t(x++) * 4 + size_t(c)] = v;
          } else {
            if (count == 0 || x + count > w) raise(NE_ERR_PARSE, path + ": bad literal run");
            for (int k = 0; k < count; ++k) scan[size_t(x++) * 4 + size_t(c)] = byte();
          }
        }
      }
    } else {
      for (size_t k = 0; k < scan.size(); ++k) scan[k] = byte();
    }
    for (int x = 0; x < w; ++x) {
      const unsigned char* q = &scan[size_t(x) * 4];
      float* o = &img.rgb[(size_t(y) * size_t(w) + size_t(x)) * 3];
      if (q[3] == 0) { o[0] = o[1] = o[2] = 0.0f; continue; }
      const float f = std::ldexp(1.0f, int(q[3]) - 136);
      o[0] = float(q[0]) * f;
      o[1] = float(q[1]) * f;
      o[2] = float(q[2]) * f;
    }
  }
}

// ---------------------------------------------------------------------------
// Instance transforms: translate * eulerAngleXYZ(rotation in degrees) * scale
// (getTransform, src/utils/Math.h:852-859), row-major, FP64.
// ---------------------------------------------------------------------------
struct Mat4 {
  double m[16];
};

Mat4 instance_transform(const double pos[3], const double rot_deg[3], const double scale[3]) {
  const double k = 3.14159265358979323846 / 180.0;
  const double cx = std::cos(rot_deg[0] * k), sx = std::sin(rot_deg[0] * k);
  const double cy = std::cos(rot_deg[1] * k), sy = std::sin(rot_deg[1] * k);
  const double cz = std::cos(rot_deg[2] * k), sz = std::sin(rot_deg[2] * k);
  const double Rx[9] = {1, 0, 0, 0, cx, -sx, 0, sx, cx};
  const double Ry[9] = {cy, 0, sy, 0, 1, 0, -sy, 0, cy};
  const double Rz[9] = {cz, -sz, 0, sz, cz, 0, 0, 0, 1};
  double Rxy[9], R[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      Rxy[3 * i + j] = Rx[3 * i] * Ry[j] + Rx[3 * i + 1] * Ry[3 + j] + Rx[3 * i + 2] * Ry[6 + j];
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      R[3 * i + j] = Rxy[3 * i] * Rz[j] + Rxy[3 * i + 1] * Rz[3 + j] + Rxy[3 * i + 2] * Rz[6 + j];
  Mat4 M{};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) M.m[4 * i + j] = R[3 * i + j] * scale[j];
    M.m[4 * i + 3] = pos[i];
  }
  M.m[15] = 1;
  return M;
}

Mat4 affine_inverse(const Mat4& M) {
  const double* a = M.m;
  const double det = a[0] * (a[5] * a[10] - a[6] * a[9]) - a[1] * (a[4] * a[10] - a[6] * a[8]) +
                     a[2] * (a[4] * a[9] - a[5] * a[8]);
  if (!(std::fabs(det) > 0)) raise(NE_ERR_PARSE, "singular instance transform (a zero scale)");
  const double id = 1.0 / det;
  double I[9] = {(a[5] * a[10] - a[6] * a[9]) * id, (a[2] * a[9] - a[1] * a[10]) * id, (a[1] * a[6] - a[2] * a[5]) * id,
                 (a[6] * a[8] - a[4] * a[10]) * id, (a[0] * a[10] - a[2] * a[8]) * id, (a[2] * a[4] - a[0] * a[6]) * id,
                 (a[4] * a[9] - a[5] * a[8]) * id, (a[1] * a[8] - a[0] * a[9]) * id, (a[0] * a[5] - a[1] * a[4]) * id};
  Mat4 R{};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) R.m[4 * i + j] = I[3 * i + j];
    R.m[4 * i + 3] = -(I[3 * i] * a[3] + I[3 * i + 1] * a[7] + I[3 * i + 2] * a[11]);
  }
  R.m[15] = 1;
  return R;
}

void xf_point(const Mat4& M, const double p[3], double out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = M.m[4 * i] * p[0] + M.m[4 * i + 1] * p[1] + M.m[4 * i + 2] * p[2] + M.m[4 * i + 3];
}
void xf_vec(const Mat4& M, const double v[3], double out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = M.m[4 * i] * v[0] + M.m[4 * i + 1] * v[1] + M.m[4 * i + 2] * v[2];
}

// ---------------------------------------------------------------------------
// Wavefront OBJ: "v x y z", "f a b c ..." with a, a/t, a//n, a/t/n indices
// (1-based, negative = relative to the end); polygons are fan-triangulated.
// ---------------------------------------------------------------------------
void read_obj(const std::string& path, std::vector<double>& verts, std::vector<uint32_t>& tris) {
  const std::string text = read_file(path);
  const char* p = text.data();
  const char* e = p + text.size();
  int line = 0;
  std::vector<int64_t> face;
  while (p < e) {
    ++line;
    const char* le = static_cast<const char*>(std::memchr(p, '\n', size_t(e - p)));
    if (!le) le = e;
    const char* hash = static_cast<const char*>(std::memchr(p, '#', size_t(le - p)));
    Tokens t{p, hash ? hash : le};
    const char *b, *q;
    auto fail_line = [&](const char* what) {
      raise(NE_ERR_PARSE, path + " line " + std::to_string(line) + ": " + what);
    };
    if (t.next(b, q)) {
      const std::string key(b, q);
      if (key == "v") {
        for (int a = 0; a < 3; ++a) {
          float v;
          if (!t.next(b, q) || !parse_float(b, q, v)) fail_line("a vertex needs three coordinates");
          verts.push_back(v);
        }
      } else if (key == "f") {
        face.clear();
        while (t.next(b, q)) {
          const char* slash = static_cast<const char*>(std::memchr(b, '/', size_t(q - b)));
          std::string tok(b, slash ? slash : q);
          char* end = nullptr;
          long long idx = std::strtoll(tok.c_str(), &end, 10);
          if (tok.empty() || end != tok.c_str() + tok.size() || idx == 0) fail_line("invalid face index");
          const int64_t nv = int64_t(verts.size() / 3);
          const int64_t k = idx > 0 ? idx - 1 : nv + idx;
          if (k < 0 || k >= nv) fail_line("face index out of range");
          face.push_back(k);
        }
        if (face.size() < 3) fail_line("a face needs at least three vertices");
        for (size_t i = 1; i + 1 < face.size(); ++i) {
          tris.push_back(uint32_t(face[0]));
          tris.push_back(uint32_t(face[i]));
          tris.push_back(uint32_t(face[i + 1]));
        }
      }
      // vt, vn, o, g, s, usemtl, mtllib, l, p: not used by the path
    }
    p = le < e ? le + 1 : e;
  }
}

// ---------------------------------------------------------------------------
// Scene flattening (SceneReader::processMaterial / processPrimitives /
// processCameraAndRenderer).
// ---------------------------------------------------------------------------
struct MaterialDef {
  enum Kind { kSurface, kSun, kEnv, kVolume } kind = kSurface;
  int32_t record = -1;        // surface: index of its 12-float record
  bool emitter = false;
  bool odd_normal_map = false;
  float sun_dir[3] = {0, 0, 0}, sun_E[3] = {0, 0, 0};
  float sig_s[3] = {0, 0, 0}, sig_a[3] = {0, 0, 0}, g = 0, density = 0;
  bool has_grid = false;
  bool grid_unavailable = false;  // lenient: a .vdb grid
  int32_t res[3] = {0, 0, 0};
  std::vector<float> grid;
  Image env;  // kEnv: the radiance map
};

}  // namespace

struct ne_scene_file {
  std::vector<float> tri_v, sph, rect, mat, grid, env, point;
  std::vector<int32_t> tri_mat, sph_mat, rect_mat, point_mat;
  ne_scene_desc desc{};
  ne_camera_desc camera{};
  ne_scene_settings settings{};
  std::string warnings;
};

namespace {

class Loader {
 public:
  Loader(ne_scene_file& out, std::string base, int flags, int maj_res)
      : f_(out), base_(std::move(base)), lenient_(flags & NE_LOAD_LENIENT), maj_res_(maj_res) {}

  void load(const Json& doc) {
    if (doc.kind != Json::kObj) raise(NE_ERR_PARSE, "Malformatted json file: not a valid json object.");
    if (!doc.find("version")) raise(NE_ERR_PARSE, "Incomplete json file: version is not present.");
    const Json& mats = member(doc, "materials", "scene");
    if (mats.kind != Json::kArr) raise(NE_ERR_PARSE, "scene: materials must be an array");
    for (const Json& m : mats.arr) material(m);
    const Json& prims = member(doc, "primitives", "scene");
    if (prims.kind != Json::kArr) raise(NE_ERR_PARSE, "scene: primitives must be an array");
    for (const Json& p : prims.arr) primitive(p);
    camera_and_renderer(member(doc, "camera", "scene"), member(doc, "renderer", "scene"));
    finish();
  }

 private:
  ne_scene_file& f_;
  std::string base_;
  bool lenient_;
  int maj_res_;
  std::map<std::string, MaterialDef> mats_;
  bool have_volume_ = false;
  const MaterialDef* volume_mat_ = nullptr;
  Mat4 vol_xf_{};

  void warn(const std::string& w) { f_.warnings += w + "\n"; }
  void unsupported(const std::string& what) { raise(NE_ERR_UNSUPPORTED, what); }

  static const Json& member(const Json& o, const char* key, const std::string& ctx) {
    const Json* v = o.find(key);
    if (!v) raise(NE_ERR_PARSE, ctx + " is missing the property " + key + ".");
    return *v;
  }
  static double number(const Json& o, const char* key, const std::string& ctx) {
    const Json& v = member(o, key, ctx);
    if (v.kind != Json::kNum) raise(NE_ERR_PARSE, ctx + ": " + key + " must be a number");
    return v.num;
  }
  static int32_t integer(const Json& v, const std::string& what) {
    if (v.kind != Json::kNum || v.num != std::floor(v.num) || std::fabs(v.num) > 2147483647.0)
      raise(NE_ERR_PARSE, what + " must be an integer");
    return int32_t(v.num);
  }
  static std::string text(const Json& o, const char* key, const std::string& ctx) {
    const Json& v = member(o, key, ctx);
    if (v.kind != Json::kStr) raise(NE_ERR_PARSE, ctx + ": " + key + " must be a string");
    return v.str;
  }
  static bool boolean(const Json& o, const char* key, const std::string& ctx) {
    const Json& v = member(o, key, ctx);
    if (v.kind != Json::kBool) raise(NE_ERR_PARSE, ctx + ": " + key + " must be true or false");
    return v.b;
  }
  // getVec3 (SceneReader.cpp:60-65): three numbers.
  static void vec3(const Json& v, double out[3], const std::string& what) {
    if (v.kind != Json::kArr || v.arr.size() != 3) raise(NE_ERR_PARSE, what + " must be an array of three numbers");
    for (int a = 0; a < 3; ++a) {
      if (v.arr[size_t(a)].kind != Json::kNum) raise(NE_ERR_PARSE, what + " must be an array of three numbers");
      out[a] = v.arr[size_t(a)].num;
    }
  }
  static void vec3(const Json& o, const char* key, double out[3], const std::string& ctx) {
    vec3(member(o, key, ctx), out, ctx + ": " + key);
  }
  int32_t add_record(int type, const double albedo[3], double rough, double metal, const double* radiance) {
    const size_t k = f_.mat.size();
    f_.mat.resize(k + 12, 0.0f);
    f_.mat[k] = float(type);
    for (int a = 0; a < 3; ++a) {
      if (albedo) f_.mat[k + 1 + size_t(a)] = float(albedo[a]);
      if (radiance) f_.mat[k + 6 + size_t(a)] = float(radiance[a]);
    }
    f_.mat[k + 4] = float(rough);
    f_.mat[k + 5] = float(metal);
    return int32_t(k / 12);
  }

  // processMaterial, SceneReader.cpp:67-222
  void material(const Json& m) {
    if (m.kind != Json::kObj || !m.find("name")) raise(NE_ERR_PARSE, "Material is missing the property name.");
    const std::string name = text(m, "name", "Material");
    const std::string ctx = "Material " + name;
    if (!m.find("type")) raise(NE_ERR_PARSE, ctx + " is missing the property type.");
    const std::string type = text(m, "type", ctx);
    MaterialDef d;
    if (type == "microfacet") {
      double albedo[3] = {1, 0, 1};
      const Json& a = member(m, "albedo", ctx);
      if (a.kind == Json::kStr) {
        if (!lenient_) unsupported(ctx + ": image albedo textures are not supported");
        albedo[0] = albedo[1] = albedo[2] = 0.5;
        warn(ctx + ": image albedo '" + a.str + "' replaced by grey 0.5");
      } else {
        vec3(a, albedo, ctx + ": albedo");
      }
      if (const Json* nm = m.find("normalMap")) {
        double n[3];
        vec3(*nm, n, ctx + ": normalMap");
        d.odd_normal_map = !(n[0] == 0.5 && n[1] == 0.5 && n[2] == 1.0);  // (0.5,0.5,1) keeps the normal
      }
      const double rough = number(m, "roughness", ctx), metal = number(m, "metallic", ctx);
      d.kind = MaterialDef::kSurface;
      d.record = add_record(NE_MAT_GLOSSY, albedo, rough, metal, nullptr);
    } else if (type == "emitter") {
      double le[3];
      vec3(m, "albedo", le, ctx);
      d.kind = MaterialDef::kSurface;
      d.emitter = true;
      d.record = add_record(NE_MAT_EMITTER, nullptr, 0.0, 0.0, le);
    } else if (type == "directionalLight") {
      double le[3], pos[3];
      vec3(m, "albedo", le, ctx);
      vec3(m, "position", pos, ctx);
      const double n = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
      if (!(n > 0)) raise(NE_ERR_PARSE, ctx + ": a directional light at the origin has no direction");
      d.kind = MaterialDef::kSun;
      for (int a = 0; a < 3; ++a) {
        d.sun_dir[a] = float(-pos[a] / n);  // normalize(vec3(0) - position)
        d.sun_E[a] = float(le[a]);
      }
    } else if (type == "infiniteAreaLight") {
      const std::string rel = text(m, "path", ctx);
      const std::string path = join_path(base_, rel);
      d.kind = MaterialDef::kEnv;
      if (ends_with(rel, ".hdr")) read_hdr(path, d.env);
      else if (ends_with(rel, ".pfm")) read_pfm(path, d.env);
      else unsupported(ctx + ": environment images must be Radiance .hdr or .pfm ('" + rel + "')");
    } else if (type == "volume") {
      double ss[3], sa[3];
      vec3(m, "scattering", ss, ctx);
      vec3(m, "absorption", sa, ctx);
      const std::string pf = text(m, "phaseFunction", ctx);
      if (pf == "isotropic") d.g = 0.0f;
      else if (pf == "hg" || pf == "henyey-greenstein") d.g = float(number(m, "g", ctx));
      else raise(NE_ERR_PARSE, ctx + ": unknown phase function '" + pf + "'");
      for (int a = 0; a < 3; ++a) { d.sig_s[a] = float(ss[a]); d.sig_a[a] = float(sa[a]); }
      d.density = float(number(m, "density", ctx));
      d.kind = MaterialDef::kVolume;
      if (const Json* pj = m.find("path")) {
        if (pj->kind != Json::kStr) raise(NE_ERR_PARSE, ctx + ": path must be a string");
        const std::string path = join_path(base_, pj->str);
        d.has_grid = true;
        if (pj->str.find(".vdb") != std::string::npos) {
          if (!lenient_) unsupported(ctx + ": OpenVDB grids are not supported (convert to .vol or .f32)");
          d.grid_unavailable = true;
          warn(ctx + ": OpenVDB grid '" + pj->str + "' cannot be read; the material is unusable");
        } else if (pj->str.find(".vol") != std::string::npos) {
          read_vol(path, d.res, &d.grid);
        } else if (ends_with(pj->str, ".f32")) {
          const Json& r = member(m, "resolution", ctx);
          if (r.kind != Json::kArr || r.arr.size() != 3) raise(NE_ERR_PARSE, ctx + ": resolution must be [W, H, D]");
          for (int a = 0; a < 3; ++a) {
            d.res[a] = integer(r.arr[size_t(a)], ctx + ": resolution");
            if (d.res[a] < 1 || d.res[a] > 4096) raise(NE_ERR_PARSE, ctx + ": resolution out of [1, 4096]");
          }
          read_f32(path, d.res, d.grid);
        } else {
          raise(NE_ERR_PARSE, ctx + ": unknown grid file type '" + pj->str + "'");
        }
      }
    } else {
      raise(NE_ERR_PARSE, "Invalid material type '" + type + "' while reading scene from json file.");
    }
    mats_[name] = std::move(d);  // replaceMaterial: the last definition of a name wins
  }

  const MaterialDef& material_of(const Json& p, const std::string& ctx) {
    const std::string name = text(p, "materialName", ctx);
    auto it = mats_.find(name);
    if (it == mats_.end()) raise(NE_ERR_PARSE, ctx + ": unknown material '" + name + "'");
    return it->second;
  }
  static void transform_of(const Json& p, const std::string& ctx, bool full, Mat4& M, double pos[3]) {
    const Json& t = member(p, "transform", ctx);
    vec3(t, "position", pos, ctx + ": transform");
    double rot[3] = {0, 0, 0}, scale[3] = {1, 1, 1};
    if (full) {
      vec3(t, "rotation", rot, ctx + ": transform");
      vec3(t, "scale", scale, ctx + ": transform");
    }
    M = instance_transform(pos, rot, scale);
  }
  void add_triangles(const std::vector<double>& v, const std::vector<uint32_t>& t, const Mat4& M, int32_t rec) {
    std::vector<double> w(v.size());
    for (size_t i = 0; i < v.size() / 3; ++i) xf_point(M, &v[3 * i], &w[3 * i]);
    for (size_t k = 0; k < t.size(); ++k)
      for (int a = 0; a < 3; ++a) f_.tri_v.push_back(float(w[3 * size_t(t[k]) + size_t(a)]));
    f_.tri_mat.insert(f_.tri_mat.end(), t.size() / 3, rec);
  }

  // processPrimitives, SceneReader.cpp:224-648
  void primitive(const Json& p) {
    if (p.kind != Json::kObj || !p.find("name")) raise(NE_ERR_PARSE, "Primitive is missing the property name.");
    const std::string name = text(p, "name", "Primitive");
    const std::string ctx = "Primitive " + name;
    if (!p.find("type")) raise(NE_ERR_PARSE, ctx + " is missing the property type.");
    const std::string type = text(p, "type", ctx);
    Mat4 M;
    double pos[3];
    if (type == "obj" || type == "gltf") {
      transform_of(p, ctx, true, M, pos);
      const std::string path = text(p, "path", ctx);
      if (type == "gltf") unsupported(ctx + ": glTF models are not supported");
      int32_t rec;
      if (p.find("materialName")) {
        const MaterialDef& d = material_of(p, ctx);
        if (d.kind != MaterialDef::kSurface) raise(NE_ERR_PARSE, ctx + ": a mesh needs a surface material");
        if (d.odd_normal_map) {
          if (!lenient_) unsupported(ctx + ": constant normal maps other than (0.5, 0.5, 1) are not supported");
          warn(ctx + ": normal map ignored");
        }
        rec = d.record;
      } else {
        if (!lenient_) unsupported(ctx + ": materials from the model file need image textures (give materialName)");
        const double grey[3] = {0.5, 0.5, 0.5};
        rec = add_record(NE_MAT_GLOSSY, grey, 0.65, 0.0, nullptr);
        warn(ctx + ": model materials replaced by glossy grey 0.5, roughness 0.65");
      }
      std::vector<double> v;
      std::vector<uint32_t> t;
      read_obj(join_path(base_, path), v, t);
      add_triangles(v, t, M, rec);
    } else if (type == "point") {
      transform_of(p, ctx, true, M, pos);
      const MaterialDef& d = material_of(p, ctx);
      if (d.kind == MaterialDef::kSun) {
        if (f_.desc.has_sun) unsupported(ctx + ": one directional light per scene");
        f_.desc.has_sun = 1;
        for (int a = 0; a < 3; ++a) { f_.desc.sun_dir[a] = d.sun_dir[a]; f_.desc.sun_E[a] = d.sun_E[a]; }
      } else if (d.kind == MaterialDef::kEnv) {
        if (f_.desc.env_type != NE_ENV_NONE) unsupported(ctx + ": one environment light per scene");
        f_.desc.env_type = NE_ENV_MAP;
        f_.env = d.env.rgb;
        f_.desc.env_w = d.env.w;
        f_.desc.env_h = d.env.h;
      } else if (d.kind == MaterialDef::kSurface && d.emitter) {
        for (int a = 0; a < 3; ++a) f_.point.push_back(float(pos[a]));
        f_.point_mat.push_back(d.record);
      }
      // a point with a non-emitting material has no area: nothing to add
    } else if (type == "sphere") {
      transform_of(p, ctx, false, M, pos);
      const double r = number(p, "radius", ctx);
      const MaterialDef& d = material_of(p, ctx);
      if (d.kind != MaterialDef::kSurface) raise(NE_ERR_PARSE, ctx + ": a sphere needs a surface material");
      bool collision = true;
      if (p.find("collision")) collision = boolean(p, "collision", ctx);
      if (!collision) {
        if (!d.emitter) return;  // invisible and not a light: it contributes nothing
        if (!lenient_) unsupported(ctx + ": non-colliding emitter spheres are not supported");
        warn(ctx + ": non-colliding emitter sphere made colliding");
      }
      if (!(r > 0)) raise(NE_ERR_PARSE, ctx + ": radius must be > 0");
      for (int a = 0; a < 3; ++a) f_.sph.push_back(float(pos[a]));
      f_.sph.push_back(float(r));
      f_.sph_mat.push_back(d.record);
    } else if (type == "rectangle") {
      transform_of(p, ctx, true, M, pos);
      const MaterialDef& d = material_of(p, ctx);
      if (d.kind != MaterialDef::kSurface) raise(NE_ERR_PARSE, ctx + ": a rectangle needs a surface material");
      // unit quad [-.5,.5]^2 at z = 0 with normal (0,0,-1), SceneReader.cpp:407-419
      const double c0[3] = {-0.5, -0.5, 0.0}, ex[3] = {1, 0, 0}, ey[3] = {0, 1, 0}, nz[3] = {0, 0, -1};
      double p0[3], eu[3], ev[3], n[3];
      xf_point(M, c0, p0);
      xf_vec(M, ex, eu);
      xf_vec(M, ey, ev);
      xf_vec(M, nz, n);
      const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      if (!(nn > 0)) raise(NE_ERR_PARSE, ctx + ": degenerate transform");
      for (double x : p0) f_.rect.push_back(float(x));
      for (double x : eu) f_.rect.push_back(float(x));
      for (double x : ev) f_.rect.push_back(float(x));
      for (double x : n) f_.rect.push_back(float(x / nn));
      f_.rect_mat.push_back(d.record);
    } else if (type == "volume") {
      transform_of(p, ctx, true, M, pos);
      const MaterialDef& d = material_of(p, ctx);
      if (d.kind != MaterialDef::kVolume) raise(NE_ERR_PARSE, ctx + ": a volume needs a volume material");
      if (d.grid_unavailable) unsupported(ctx + ": its material's grid could not be read");
      if (have_volume_) unsupported(ctx + ": one volume per scene");
      have_volume_ = true;
      volume_mat_ = &d;
      vol_xf_ = M;
    } else {
      raise(NE_ERR_PARSE, "Invalid primitive type '" + type + "' while reading scene from json file.");
    }
  }

  // processCameraAndRenderer, SceneReader.cpp:597-621
  void camera_and_renderer(const Json& cam, const Json& ren) {
    double position[3], look_at[3], up[3];
    vec3(cam, "position", position, "camera");
    vec3(cam, "lookAt", look_at, "camera");
    vec3(cam, "up", up, "camera");
    number(cam, "speed", "camera");
    const double vfov = number(cam, "vfov", "camera");
    number(cam, "aperture", "camera");
    bool auto_focus = false;
    if (cam.find("autoFocus")) auto_focus = boolean(cam, "autoFocus", "camera");
    if (!auto_focus) number(cam, "focus", "camera");
    const Json& r = member(ren, "resolution", "renderer");
    if (r.kind != Json::kArr || r.arr.size() != 2) raise(NE_ERR_PARSE, "renderer: resolution must be [width, height]");
    const int32_t w = integer(r.arr[0], "renderer: resolution"), h = integer(r.arr[1], "renderer: resolution");
    if (w < 1 || h < 1) raise(NE_ERR_PARSE, "renderer: resolution must be positive");
    f_.settings.width = w;
    f_.settings.height = h;
    f_.settings.spp = integer(member(ren, "spp", "renderer"), "renderer: spp");
    f_.settings.bounces = integer(member(ren, "bounces", "renderer"), "renderer: bounces");
    f_.settings.hdr = boolean(ren, "HDR", "renderer") ? 1 : 0;
    f_.settings.realtime = text(ren, "mode", "renderer") == "realtime" ? 1 : 0;
    // Camera(position, lookAt, vec3(0,1,0), vfov, width / height, ...): the file's "up" is not used
    for (int a = 0; a < 3; ++a) {
      f_.camera.position[a] = float(position[a]);
      f_.camera.look_at[a] = float(look_at[a]);
    }
    f_.camera.up[0] = 0.0f;
    f_.camera.up[1] = 1.0f;
    f_.camera.up[2] = 0.0f;
    f_.camera.vfov = float(vfov);
    f_.camera.aspect = float(w) / float(h);
  }

  void finish() {
    ne_scene_desc& d = f_.desc;
    if (have_volume_) {
      const MaterialDef& m = *volume_mat_;
      d.has_volume = 1;
      const Mat4 inv = affine_inverse(vol_xf_);
      for (int i = 0; i < 16; ++i) { d.vol_xf[i] = float(vol_xf_.m[i]); d.vol_inv[i] = float(inv.m[i]); }
      d.vol_mult = m.density;
      d.vol_g = m.g;
      for (int a = 0; a < 3; ++a) { d.vol_sigma_s[a] = m.sig_s[a]; d.vol_sigma_a[a] = m.sig_a[a]; }
      if (m.has_grid) {
        f_.grid = m.grid;
        for (int a = 0; a < 3; ++a) d.vol_res[a] = m.res[a];
        const int mx = std::max(m.res[0], std::max(m.res[1], m.res[2]));
        d.maj_res = maj_res_ > 0 ? maj_res_ : std::max(1, std::min(16, mx));
      } else {
        d.vol_homogeneous = 1;
        d.maj_res = 1;
      }
    }
  }
};

void point_desc(ne_scene_file* f) {
  ne_scene_desc& d = f->desc;
  d.tri_v = f->tri_v.empty() ? nullptr : f->tri_v.data();
  d.tri_mat = f->tri_mat.empty() ? nullptr : f->tri_mat.data();
  d.n_tri = uint32_t(f->tri_mat.size());
  d.sph = f->sph.empty() ? nullptr : f->sph.data();
  d.sph_mat = f->sph_mat.empty() ? nullptr : f->sph_mat.data();
  d.n_sph = uint32_t(f->sph_mat.size());
  d.rect = f->rect.empty() ? nullptr : f->rect.data();
  d.rect_mat = f->rect_mat.empty() ? nullptr : f->rect_mat.data();
  d.n_rect = uint32_t(f->rect_mat.size());
  d.mat = f->mat.empty() ? nullptr : f->mat.data();
  d.n_mat = uint32_t(f->mat.size() / 12);
  d.grid = f->grid.empty() ? nullptr : f->grid.data();
  d.env_map = f->env.empty() ? nullptr : f->env.data();
  d.point = f->point.empty() ? nullptr : f->point.data();
  d.point_mat = f->point_mat.empty() ? nullptr : f->point_mat.data();
  d.n_point = uint32_t(f->point_mat.size());
}

ne_status parse_into(const char* json, size_t len, const ne_load_options* opt, const std::string& default_base,
                     ne_scene_file** out) {
  if (!out) return ne::host_fail(NE_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (!json) return ne::host_fail(NE_ERR_INVALID_ARG, "json is NULL");
  const int flags = opt ? opt->flags : 0;
  const int maj = opt ? opt->maj_res : 0;
  if (maj < 0 || maj > 256) return ne::host_fail(NE_ERR_INVALID_ARG, "maj_res must be in [0, 256]");
  const std::string base = opt && opt->resources_dir ? std::string(opt->resources_dir) : default_base;
  ne_scene_file* f = new ne_scene_file();
  try {
    JsonParser jp(json, len);
    const Json doc = jp.parse();
    Loader(*f, base, flags, maj).load(doc);
  } catch (const Failure& e) {
    delete f;
    return ne::host_fail(e.code, e.msg);
  } catch (const std::bad_alloc&) {
    delete f;
    return ne::host_fail(NE_ERR_OUT_OF_MEMORY, "out of host memory while reading the scene");
  }
  point_desc(f);
  *out = f;
  return NE_OK;
}

}  // namespace

extern "C" {

ne_status ne_scene_file_parse(const char* json, size_t length, const ne_load_options* opt, ne_scene_file** out) {
  return parse_into(json, length, opt, std::string(), out);
}

ne_status ne_scene_file_read(const char* path, const ne_load_options* opt, ne_scene_file** out) {
  if (out) *out = nullptr;
  if (!path) return ne::host_fail(NE_ERR_INVALID_ARG, "path is NULL");
  std::string text;
  try {
    text = read_file(path);
  } catch (const Failure& e) {
    return ne::host_fail(e.code, e.msg);
  }
  return parse_into(text.data(), text.size(), opt, dir_of(path), out);
}

ne_status ne_scene_file_get(const ne_scene_file* f, ne_scene_desc* desc, ne_camera_desc* camera,
                            ne_scene_settings* settings) {
  if (!f) return ne::host_fail(NE_ERR_INVALID_ARG, "scene file is NULL");
  if (desc) *desc = f->desc;
  if (camera) *camera = f->camera;
  if (settings) *settings = f->settings;
  return NE_OK;
}

const char* ne_scene_file_warnings(const ne_scene_file* f) { return f ? f->warnings.c_str() : ""; }

void ne_scene_file_release(ne_scene_file* f) { delete f; }

ne_status ne_vol_read(const char* path, int32_t res[3], float* grid, uint64_t capacity) {
  if (!path || !res) return ne::host_fail(NE_ERR_INVALID_ARG, "path and res must not be NULL");
  try {
    if (!grid) {
      read_vol(path, res, nullptr);
      return NE_OK;
    }
    std::vector<float> g;
    int32_t r[3];
    read_vol(path, r, &g);
    if (capacity < g.size()) return ne::host_fail(NE_ERR_INVALID_ARG, "grid capacity below W*H*D");
    std::memcpy(grid, g.data(), g.size() * sizeof(float));
    for (int a = 0; a < 3; ++a) res[a] = r[a];
  } catch (const Failure& e) {
    return ne::host_fail(e.code, e.msg);
  } catch (const std::bad_alloc&) {
    return ne::host_fail(NE_ERR_OUT_OF_MEMORY, "out of host memory while reading the grid");
  }
  return NE_OK;
}

}  // extern "C"
