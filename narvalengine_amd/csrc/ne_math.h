// ne_math.h -- host/device vector math and the counter-based RNG used by
// every kernel of the path-tracing backend.
//
// Conventions follow the reference's glm usage (GLSL semantics):
//   normalize(v) = v * (1 / sqrt(dot(v, v))), mix(x, y, a) = x*(1-a) + y*a,
//   cross in glm's operand order.  All code is built with -ffp-contract=off
//   so each multiply and add rounds on its own, which keeps device results
//   bit-comparable with a plain C++ build of the same formulas.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NE_HD __host__ __device__ __forceinline__
#define NE_DEV __device__ __forceinline__

namespace ne {

constexpr double kPI = 3.14159265358979323846264338327950288;
constexpr double kTWO_PI = 6.283185307179586476925286766559;
constexpr double kFOUR_PI = 12.566370614359172953850573533118;
constexpr float kINV4PI = float(0.07957747154594766788);
constexpr double kEPS = 0.0000000001;  // EPSILON, src/utils/Math.h:20

// ---- scalar overloads (float / double) ------------------------------------
NE_HD float xsqrt(float x) { return sqrtf(x); }
NE_HD double xsqrt(double x) { return sqrt(x); }
NE_HD float xlog(float x) { return logf(x); }
NE_HD double xlog(double x) { return log(x); }
NE_HD float xatan2(float y, float x) { return atan2f(y, x); }
NE_HD double xatan2(double y, double x) { return atan2(y, x); }
NE_HD float xacos(float x) { return acosf(x); }
NE_HD double xacos(double x) { return acos(x); }
NE_HD float xexp(float x) { return expf(x); }
NE_HD double xexp(double x) { return exp(x); }
NE_HD float xsin(float x) { return sinf(x); }
NE_HD double xsin(double x) { return sin(x); }
NE_HD float xcos(float x) { return cosf(x); }
// sin and cos with one argument reduction (the same values as xsin, xcos)
NE_HD void xsincos(float x, float& s, float& c) { sincosf(x, &s, &c); }
NE_HD void xsincos(double x, double& s, double& c) { sincos(x, &s, &c); }
NE_HD double xcos(double x) { return cos(x); }
NE_HD float xfabs(float x) { return fabsf(x); }
NE_HD double xfabs(double x) { return fabs(x); }
NE_HD float xfloor(float x) { return floorf(x); }
NE_HD double xfloor(double x) { return floor(x); }
template <typename R> NE_HD R xmax(R a, R b) { return (a < b) ? b : a; }  // std::max
template <typename R> NE_HD R xmin(R a, R b) { return (b < a) ? b : a; }  // std::min
template <typename R> NE_HD R xinf();
template <> NE_HD float xinf<float>() { return __builtin_huge_valf(); }
template <> NE_HD double xinf<double>() { return __builtin_huge_val(); }
NE_HD bool xisinf(float x) { return __builtin_isinf(x); }
NE_HD bool xisinf(double x) { return __builtin_isinf(x); }

// ---- vec3 -----------------------------------------------------------------
template <typename R>
struct Vec3 {
  R x, y, z;
};
template <typename R> NE_HD Vec3<R> v3(R a, R b, R c) { return Vec3<R>{a, b, c}; }
template <typename R> NE_HD Vec3<R> v3s(R a) { return Vec3<R>{a, a, a}; }
template <typename R> NE_HD Vec3<R> v3f(const float* p) { return Vec3<R>{R(p[0]), R(p[1]), R(p[2])}; }
template <typename R> NE_HD Vec3<R> v3f4(float4 p) { return Vec3<R>{R(p.x), R(p.y), R(p.z)}; }
template <typename R, typename S> NE_HD Vec3<R> vcast(Vec3<S> a) { return Vec3<R>{R(a.x), R(a.y), R(a.z)}; }
template <typename R> NE_HD Vec3<R> operator+(Vec3<R> a, Vec3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> NE_HD Vec3<R> operator-(Vec3<R> a, Vec3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> NE_HD Vec3<R> operator*(Vec3<R> a, Vec3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename R> NE_HD Vec3<R> operator/(Vec3<R> a, Vec3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <typename R> NE_HD Vec3<R> operator*(Vec3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R> NE_HD Vec3<R> operator*(R s, Vec3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename R> NE_HD Vec3<R> operator/(Vec3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename R> NE_HD Vec3<R> operator-(Vec3<R> a) { return {-a.x, -a.y, -a.z}; }
template <typename R> NE_HD R dot(Vec3<R> a, Vec3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> NE_HD Vec3<R> cross(Vec3<R> a, Vec3<R> b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <typename R> NE_HD R length(Vec3<R> a) { return xsqrt(dot(a, a)); }
template <typename R> NE_HD Vec3<R> normalize(Vec3<R> a) { return a * (R(1) / xsqrt(dot(a, a))); }
template <typename R> NE_HD R avg3(Vec3<R> v) { return (v.x + v.y + v.z) / R(3); }
template <typename R> NE_HD bool is_black(Vec3<R> v) { return v.x == 0 && v.y == 0 && v.z == 0; }
template <typename R> NE_HD R comp(Vec3<R> v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
NE_HD float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }

// generateOrthonormalCS (src/utils/Math.h:591-599): v, u span the plane of n.
template <typename R>
NE_HD void onb(Vec3<R> n, Vec3<R>& v, Vec3<R>& u) {
  if (xfabs(n.x) > xfabs(n.y))
    v = v3(-n.z, R(0), n.x) / xsqrt(n.x * n.x + n.z * n.z);
  else
    v = v3(R(0), n.z, -n.y) / xsqrt(n.y * n.y + n.z * n.z);
  u = normalize(cross(n, v));
}
// toWorld / toLCS (src/utils/Math.h:612-631)
template <typename R>
NE_HD Vec3<R> to_world(Vec3<R> v, Vec3<R> ns, Vec3<R> ss, Vec3<R> ts) {
  return v3(ss.x * v.x + ts.x * v.y + ns.x * v.z, ss.y * v.x + ts.y * v.y + ns.y * v.z,
            ss.z * v.x + ts.z * v.y + ns.z * v.z);
}
template <typename R>
NE_HD Vec3<R> to_lcs(Vec3<R> v, Vec3<R> ns, Vec3<R> ss, Vec3<R> ts) {
  return v3(dot(v, ss), dot(v, ts), dot(v, ns));
}
// powerHeuristic (src/utils/Math.h:751-753)
template <typename R> NE_HD R power_heuristic(R a, R b) { return (a * a) / (a * a + b * b); }

// ---- Philox-4x32-10 (Salmon et al., SC'11) ---------------------------------

NE_HD void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    // one 32x32->64 product per multiplier (a single v_mad_u64_u32 on gfx950
    // instead of separate high and low multiplies)
    const uint64_t p0 = uint64_t(0xD2511F53u) * c[0], p1 = uint64_t(0xCD9E8D57u) * c[2];
    const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}

// Per-path uniform stream.  Counter words: (pixel, sample, block, job);
// key: the 64-bit seed.  One Philox call yields four 32-bit draws; a uniform
// keeps the top 24 bits: u = (x >> 8) * 2^-24 in [0, 1).
// Block numbers [0, 2^31) carry the path's main stream; the shadow ray of
// scatter vertex v draws from its own sub-stream starting at block
// 2^31 + v * 2^20, so shadow rays can be traced in any order or stage.
constexpr uint32_t kShadowBlockBase = 0x80000000u;
constexpr int kMaxVertices = 2047;

struct PathRng {
  uint32_t pixel, sample, block, job, k0, k1;
  uint32_t b0, b1, b2, b3;  // undrawn words of the current block, next first
  int idx;                  // words of the current block drawn (4: none buffered)
  NE_HD void init(uint64_t seed, uint32_t px, uint32_t smp, uint32_t jb, uint32_t block0 = 0) {
    pixel = px; sample = smp; block = block0; job = jb;
    k0 = uint32_t(seed); k1 = uint32_t(seed >> 32);
    idx = 4;
  }
  // Resume after n draws of the stream that starts at block0.
  NE_HD void seek(uint32_t n, uint32_t block0 = 0) {
    block = block0 + (n >> 2);
    idx = 4;
    if (n & 3u) {
      refill();
      for (uint32_t k = 0; k < (n & 3u); ++k) shift();
    }
  }
  // Draws consumed so far (main stream).
  NE_HD uint32_t consumed() const { return (block << 2) - (idx == 4 ? 0u : uint32_t(4 - idx)); }
  NE_HD PathRng shadow_stream(int vertex) const {
    PathRng r = *this;
    r.block = kShadowBlockBase + (uint32_t(vertex) << 20);
    r.idx = 4;
    return r;
  }
  NE_HD void refill() {
    uint32_t c[4] = {pixel, sample, block, job};
    philox10(c, k0, k1);
    b0 = c[0]; b1 = c[1]; b2 = c[2]; b3 = c[3];
    block += 1;
    idx = 0;
  }
  // Drops the next buffered word.  The buffer is a queue of scalars (never an
  // indexed array) so it stays in registers.
  NE_HD void shift() {
    b0 = b1; b1 = b2; b2 = b3;
    idx += 1;
  }
  NE_HD uint32_t next_u32() {
    if (idx == 4) refill();
    const uint32_t r = b0;
    shift();
    return r;
  }
  NE_HD float next() { return float(next_u32() >> 8) * (1.0f / 16777216.0f); }
  // The same draw, but it also leaves at least one word buffered (a block
  // is generated when one word or none is left, the leftover word drawn
  // first), so that next_buffered() right after it needs no refill.  A loop
  // that draws one or two words per pass then has one refill site instead
  // of two, each of which costs the whole wave a Philox block whenever any
  // of its lanes needs one.  Same words in the same order as next().
  NE_HD float next_keep1() {
    uint32_t r;
    if (idx >= 3) {
      uint32_t c[4] = {pixel, sample, block, job};
      philox10(c, k0, k1);
      block += 1;
      if (idx == 3) { r = b0; b0 = c[0]; b1 = c[1]; b2 = c[2]; b3 = c[3]; idx = 0; }
      else { r = c[0]; b0 = c[1]; b1 = c[2]; b2 = c[3]; idx = 1; }
    } else {
      r = b0;
      shift();
    }
    return float(r >> 8) * (1.0f / 16777216.0f);
  }
  // The next buffered word; only directly after next_keep1().
  NE_HD float next_buffered() {
    const uint32_t r = b0;
    shift();
    return float(r >> 8) * (1.0f / 16777216.0f);
  }
};

}  // namespace ne
