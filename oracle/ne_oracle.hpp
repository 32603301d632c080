// ne_oracle.hpp -- CPU restatement of NarvalEngine's path-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  This header is the oracle: tests/, the smoke
// check and bench.py's cpu_baseline leg use it as the checker (and as the
// timed CPU baseline).  Product code never includes, links or calls it.
//
// Two layers:
//   1. "refcompat" restatements of individual reference functions, FP32,
//      same formulas and operation order as the reference (each cites the
//      reference file:line it follows).  These carry the per-primitive
//      parity contract (hit/miss, t, pdf values).
//   2. A full path integrator Li<T> (T = float or double) in the
//      "corrected" semantics: the reference's estimator structure
//      (VolumetricPathIntegrator::Li, src/integrators/VolumetricPathIntegrator.cpp:176-301)
//      with its physically inconsistent steps fixed (phase sampling around
//      the propagation direction, NEE at the scatter point, one-sided MIS
//      bookkeeping done PBRT-style, cosine applied once, transmittance through
//      media on shadow rays).  FP64 instantiation = the converged oracle.
//
// Random numbers: Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11), keyed per path, so a
// (pixel, sample) pair always sees the same uniform stream.
//
// Built with -ffp-contract=off so every product operation is rounded
// individually (the GPU build uses the same setting).
#pragma once
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

namespace neo {

// ---------------------------------------------------------------------------
// constants (src/utils/Math.h:18-29)
// ---------------------------------------------------------------------------
constexpr double kEPS3 = 0.001;
constexpr double kEPS = 0.0000000001;     // EPSILON
constexpr double kEPS12 = 0.00000000001;  // EPSILON12
constexpr double kPI = 3.14159265358979323846264338327950288;
constexpr double kTWO_PI = 6.283185307179586476925286766559;
constexpr double kFOUR_PI = 12.566370614359172953850573533118;
constexpr float kINV4PI = float(0.07957747154594766788);
constexpr float kINF = std::numeric_limits<float>::infinity();

// ---------------------------------------------------------------------------
// small vector type (glm::vec3 stand-in; GLSL-spec semantics)
// ---------------------------------------------------------------------------
template <class T>
struct V3 {
  T x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(T a, T b, T c) : x(a), y(b), z(c) {}
  explicit V3(T a) : x(a), y(a), z(a) {}
  template <class U>
  explicit V3(const V3<U>& o) : x(T(o.x)), y(T(o.y)), z(T(o.z)) {}
  T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T> operator/(V3<T> a, V3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
// GLSL normalize = v * inversesqrt(dot(v, v))
template <class T> inline V3<T> normalize(V3<T> a) { return a * (T(1) / std::sqrt(dot(a, a))); }
template <class T> inline V3<T> vmin(V3<T> a, V3<T> b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
template <class T> inline bool is_black(V3<T> v) { return v.x == 0 && v.y == 0 && v.z == 0; }  // Math.h:641-647
template <class T> inline bool all_one(V3<T> v) { return v.x == 1 && v.y == 1 && v.z == 1; }   // Math.h:649-651
template <class T> inline T avg(V3<T> v) { return (v.x + v.y + v.z) / T(3); }                 // Math.h:930-932
using V3f = V3<float>;

// GLSL mix(x, y, a) = x * (1 - a) + y * a (glm::lerp forwards to mix)
inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }

// ---------------------------------------------------------------------------
// Philox-4x32-10 counter-based RNG
// ---------------------------------------------------------------------------
inline void philox4x32_10(const uint32_t in_ctr[4], const uint32_t in_key[2], uint32_t out[4]) {
  uint32_t c0 = in_ctr[0], c1 = in_ctr[1], c2 = in_ctr[2], c3 = in_ctr[3];
  uint32_t k0 = in_key[0], k1 = in_key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = uint64_t(0xD2511F53u) * c0;
    uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
    uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// One path's uniform stream: counter = (pixel, sample, block, job), key = seed.
// Uniform = top 24 bits * 2^-24, in [0, 1).
struct Rng {
  uint32_t ctr[4];
  uint32_t key[2];
  uint32_t buf[4];
  int idx;
  Rng(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t job) {
    ctr[0] = pixel; ctr[1] = sample; ctr[2] = 0; ctr[3] = job;
    key[0] = uint32_t(seed); key[1] = uint32_t(seed >> 32);
    idx = 4;
  }
  uint32_t next_u32() {
    if (idx == 4) { philox4x32_10(ctr, key, buf); ctr[2] += 1; idx = 0; }
    return buf[idx++];
  }
  float next() { return float(next_u32() >> 8) * (1.0f / 16777216.0f); }
  // Shadow ray of scatter vertex v: same (pixel, sample, job, key), blocks
  // from 2^31 + v * 2^20 on, disjoint from the main stream's blocks.
  Rng shadow_stream(int vertex) const {
    Rng r = *this;
    r.ctr[2] = 0x80000000u + (uint32_t(vertex) << 20);
    r.idx = 4;
    return r;
  }
};

// ---------------------------------------------------------------------------
// 1. refcompat restatements of single reference functions (FP32)
// ---------------------------------------------------------------------------
struct Hit {
  bool hit = false;
  float tNear = kINF, tFar = -kINF;
  V3f point, normal;
};

// AABB::intersect, src/primitives/AABB.cpp:48-79 (centre/half-size slab test).
inline Hit ref_aabb_intersect(V3f bmin, V3f bmax, V3f o, V3f d) {
  Hit h;
  V3f s((d.x < 0.0) ? 1.0f : -1.0f, (d.y < 0.0) ? 1.0f : -1.0f, (d.z < 0.0) ? 1.0f : -1.0f);
  V3f invD = V3f(1.0f) / d;
  V3f size = bmax - bmin;                       // getSize (AABB.cpp:34-37)
  V3f half = size / 2.0f;                       // getHalfSize (:30-32)
  V3f center = bmax - size / 2.0f;              // getCenter (:39-42)
  V3f t1 = (center - o + s * half) * invD;
  V3f t2 = (center - o - s * half) * invD;
  V3f tmin = vmin(t1, t2), tmax = vmax(t1, t2);
  float tNear = std::max(std::max(tmin.x, tmin.y), tmin.z);
  float tFar = std::min(std::min(tmax.x, tmax.y), tmax.z);
  h.tNear = tNear; h.tFar = tFar;
  h.point = o + tNear * d;
  if (tNear > tFar) { h.tNear = kINF; h.tFar = -kINF; h.hit = false; return h; }
  h.hit = true;
  return h;
}

// Triangle::intersect, src/primitives/Triangle.cpp:49-81 (no normal map).
inline Hit ref_triangle_intersect(V3f v0, V3f v1, V3f v2, V3f o, V3f d) {
  Hit h;
  V3f v1v0 = v1 - v0, v2v0 = v2 - v0, rov0 = o - v0;
  V3f n = cross(v1v0, v2v0);
  V3f q = cross(rov0, d);
  float denom = float(1.0 / double(dot(d, n)));   // `1.0 / float` promotes to double
  float u = dot(-q, v2v0) * denom;
  float v = dot(q, v1v0) * denom;
  float t = dot(-n, rov0) * denom;
  if (std::isnan(t) || t < 0) return h;
  if (u < 0.0f || u > 1.0f || v < 0.0f || (u + v) > 1.0f) return h;
  h.hit = true; h.tNear = t; h.tFar = t; h.point = o + t * d; h.normal = normalize(n);
  return h;
}

// Sphere::intersect, src/primitives/Sphere.cpp:21-44 (accepts negative roots).
inline Hit ref_sphere_intersect(V3f c, float radius, V3f o, V3f d) {
  Hit h;
  V3f oc = o - c;
  float a = dot(d, d), b = dot(oc, d), cc = dot(oc, oc) - radius * radius;
  float disc = b * b - a * cc;
  if (disc >= 0) {
    float t1 = (-b - std::sqrt(disc)) / a;
    float t2 = (-b + std::sqrt(disc)) / a;
    h.tNear = std::min(t1, t2); h.tFar = std::max(t1, t2);
    h.point = o + h.tNear * d;
    h.normal = normalize((h.point - c) / radius);
    h.hit = true;
  }
  return h;
}

// Rectangle::getSize / containsPoint / intersect, src/primitives/Rectangle.cpp:16-76.
// vA = vertexData[0], vB = vertexData[1] (SceneReader.cpp:516-517 sets them to
// the quad's (-.5,-.5,0) and (.5,.5,0) corners).
inline V3f ref_rect_size(V3f v0, V3f v1) {
  for (int i = 0; i < 3; ++i)
    if (v0[i] == v1[i]) { v0[i] = 0; v1[i] = 0; }
  return V3f(std::fabs(v0.x) + std::fabs(v1.x), std::fabs(v0.y) + std::fabs(v1.y), std::fabs(v0.z) + std::fabs(v1.z));
}
inline bool ref_rect_contains(V3f vA, V3f vB, V3f p) {
  V3f mn = vmin(vA, vB), mx = vmax(vA, vB), size = ref_rect_size(vA, vB);
  if (size.x > 0 && (p.x < mn.x || p.x > mx.x)) return false;
  if (size.y > 0 && (p.y < mn.y || p.y > mx.y)) return false;
  if (size.z > 0 && (p.z < mn.z || p.z > mx.z)) return false;
  return true;
}
inline Hit ref_rect_intersect(V3f vA, V3f vB, V3f normal, V3f o, V3f d) {
  Hit h;
  float denom = dot(normal, d);
  if (std::fabs(denom) < float(kEPS)) return h;
  float num = dot(normal, vB - o);
  float t = num / denom;
  if (t < 0) return h;
  V3f p = o + t * d;
  if (!ref_rect_contains(vA, vB, p)) return h;
  h.hit = true; h.tNear = t; h.tFar = t; h.normal = normal; h.point = p;
  return h;
}

// Triangle::calculateArea (Triangle.cpp:149-157) and barycentricCoordinates (:97-112).
inline float ref_triangle_area(V3f a, V3f b, V3f c) { return length(cross(b - a, c - a)) / 2.0f; }
inline V3f ref_barycentric(V3f p, V3f a, V3f b, V3f c) {
  V3f v0 = b - a, v1 = c - a, v2 = p - a;
  float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
  float denom = d00 * d11 - d01 * d01;
  V3f uvw;
  uvw.y = (d11 * d20 - d01 * d21) / denom;
  uvw.z = (d00 * d21 - d01 * d20) / denom;
  uvw.x = 1.0f - uvw.y - uvw.z;
  return uvw;
}
// Triangle::contains (Triangle.cpp:83-95).
inline bool ref_triangle_contains(V3f a, V3f b, V3f c, V3f p) {
  V3f n = cross(b - a, c - a);
  float lenSq = length(n);
  float dd = dot(n, b - p);
  if (dd * dd > float(kEPS3) * lenSq) return false;
  V3f uvw = ref_barycentric(p, a, b, c);
  return uvw.x >= 0 && uvw.y >= 0 && uvw.x + uvw.y <= 1;
}

// generateOrthonormalCS / toWorld / toLCS, src/utils/Math.h:591-631.
template <class T>
inline void ref_onb(V3<T> n, V3<T>& v, V3<T>& u) {
  if (std::fabs(n.x) > std::fabs(n.y))
    v = V3<T>(-n.z, T(0), n.x) / std::sqrt(n.x * n.x + n.z * n.z);
  else
    v = V3<T>(T(0), n.z, -n.y) / std::sqrt(n.y * n.y + n.z * n.z);
  u = normalize(cross(n, v));
}
template <class T>
inline V3<T> to_world(V3<T> v, V3<T> ns, V3<T> ss, V3<T> ts) {
  return V3<T>(ss.x * v.x + ts.x * v.y + ns.x * v.z, ss.y * v.x + ts.y * v.y + ns.y * v.z,
               ss.z * v.x + ts.z * v.y + ns.z * v.z);
}
template <class T>
inline V3<T> to_lcs(V3<T> v, V3<T> ns, V3<T> ss, V3<T> ts) { return V3<T>(dot(v, ss), dot(v, ts), dot(v, ns)); }

// sampleUnitSphere(e1, e2), src/utils/Math.h:442-457.
inline V3f ref_sample_unit_sphere(float e1, float e2) {
  float theta = float(kTWO_PI * e1);
  float phi = std::acos(1.0f - 2.0f * e2);
  return V3f(std::sin(phi) * std::cos(theta), std::sin(phi) * std::sin(theta), std::cos(phi));
}

// powerHeuristic, src/utils/Math.h:751-753.
template <class T> inline T power_heuristic(T a, T b) { return (a * a) / (a * a + b * b); }

// HG::pdf / HG::eval, src/materials/Medium.h:116-128:
//   (1-g^2) / (4 pi (1 + g^2 - 2 g cos)^1.5), cos = dot(normalize(a), normalize(b)).
inline float ref_hg_eval(float g, V3f a, V3f b) {
  float cosTheta = dot(normalize(a), normalize(b));
  float denom = 1.0f + g * g - 2.0f * g * cosTheta;
  return kINV4PI * (1.0f - g * g) / (denom * std::sqrt(denom));
}
// HG::sample in its local (z-up) frame, src/materials/Medium.h:96-114.
// u0, u1 are the two uniforms the reference draws.
inline V3f ref_hg_sample_local(float g, float u0, float u1) {
  float cosTheta;
  if (std::fabs(g) < 1e-3f) {
    cosTheta = 1.0f - 2.0f * u0;
  } else {
    float sqr = (1.0f - g * g) / (1.0f + g - 2.0f * g * u0);
    sqr = sqr * sqr;
    cosTheta = -1.0f / (2.0f * g) * (1.0f + g * g - sqr);
  }
  float sinTheta = std::sqrt(1.0f - cosTheta * cosTheta);
  float phi = float(2.0f * kPI * u1);
  return V3f(std::cos(phi) * sinTheta, std::sin(phi) * sinTheta, cosTheta);  // Math.h:633-635
}

// GGXDistribution, src/core/Microfacet.cpp:5-65.
inline float ref_ggx_D(float alpha, V3f h) {
  float a2 = alpha * alpha;
  float NdotH = h.z, NdotH2 = NdotH * NdotH;
  float denom = (NdotH2 * (a2 - 1.0f) + 1.0f);
  denom = std::max(float(kEPS), float(kPI * denom * denom));
  return a2 / denom;
}
inline float ref_ggx_G1(float alpha, float NdotV) {
  float k = alpha / 2.0f;
  float denom = float(std::max(kEPS, double(NdotV) * (1.0 - k) + k));
  return NdotV / denom;
}
inline float ref_ggx_G(float alpha, V3f wi, V3f wo) {
  V3f n = normalize(wi + wo);
  float NdotV = std::max(dot(n, wi), 0.0f), NdotL = std::max(dot(n, wo), 0.0f);
  return ref_ggx_G1(alpha, NdotV) * ref_ggx_G1(alpha, NdotL);
}
inline V3f ref_ggx_sample_h(float alpha, float r0, float r1) {
  float a2 = alpha * alpha;
  float theta = std::acos(std::min(1.0f, std::sqrt((1.0f - r0) / std::max(float(kEPS), (r0 * (a2 - 1.0f) + 1.0f)))));
  float phi = float(kTWO_PI * r1);
  return V3f(std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta));
}
inline float ref_ggx_pdf(float alpha, V3f scattered, V3f h) {
  float NdotH = h.z;
  float VdotH = dot(scattered, h);
  return ref_ggx_D(alpha, h) * NdotH / (4 * VdotH);
}
// FresnelSchilck::eval, src/core/BSDF.h:150-157 (f0 = 0.04).
inline float ref_fresnel(float cosTheta) { return 0.04f + (1.0f - 0.04f) * std::pow(1.0f - cosTheta, 5.0f); }
// GlossyBSDF::eval, src/core/GlossyBSDF.cpp:26-48 (alpha = roughness^2; includes NdotL).
inline V3f ref_glossy_eval(float roughness, float metallic, V3f albedo, V3f wi, V3f wo) {
  float alpha = roughness * roughness;
  V3f H = normalize(wi + wo);
  float HdotV = dot(wi, H), NdotV = wi.z, NdotL = wo.z;
  float D = ref_ggx_D(alpha, H), G = ref_ggx_G(alpha, wi, wo), F = ref_fresnel(HdotV);
  float num = D * G * F;
  float den = 4.0f * std::max(NdotV, 0.0f) * std::max(NdotL, 0.0f);
  float spec = num / std::max(den, float(kEPS));
  float kD = (1.0f - F) * (1.0f - metallic);
  return (albedo / float(kPI) * kD + V3f(spec)) * NdotL;
}

// GridMedia::density / interpolatedDensity, src/materials/GridMedia.cpp:15-38.
struct Grid {
  int W = 0, H = 0, D = 0;
  const float* data = nullptr;
  float at(int x, int y, int z) const {  // density(): index >= res reads 0
    if (x >= W || y >= H || z >= D) return 0.0f;
    return data[size_t(z) * size_t(W) * size_t(H) + size_t(y) * size_t(W) + size_t(x)];
  }
};
inline float ref_interpolated_density(const Grid& g, V3f gp) {
  gp = vmin(vmax(V3f(0.0f), gp), V3f(float(g.W), float(g.H), float(g.D)));
  int ix = int(std::floor(gp.x)), iy = int(std::floor(gp.y)), iz = int(std::floor(gp.z));
  V3f d(gp.x - float(ix), gp.y - float(iy), gp.z - float(iz));
  float d00 = mixf(g.at(ix, iy, iz), g.at(ix + 1, iy, iz), d.x);
  float d10 = mixf(g.at(ix, iy + 1, iz), g.at(ix + 1, iy + 1, iz), d.x);
  float d01 = mixf(g.at(ix, iy, iz + 1), g.at(ix + 1, iy, iz + 1), d.x);
  float d11 = mixf(g.at(ix, iy + 1, iz + 1), g.at(ix + 1, iy + 1, iz + 1), d.x);
  float d0 = mixf(d00, d10, d.y), d1 = mixf(d01, d11, d.y);
  return mixf(d0, d1, d.z);
}
// fromOCStoGCS (GridMedia.h:33-35) + sampleAt (GridMedia.cpp:40-43).
inline float ref_density_ocs(const Grid& g, V3f p) {
  V3f gp = (p + V3f(0.5f)) * V3f(float(g.W), float(g.H), float(g.D));
  return ref_interpolated_density(g, gp);
}
// calculateMaxDensity, src/materials/GridMedia.h:16-21.
inline float ref_max_density(const Grid& g) {
  float m = 0;
  size_t n = size_t(g.W) * g.H * g.D;
  for (size_t i = 0; i < n; ++i) m = std::max(m, g.data[i]);
  return m;
}

// GridMedia::sample (delta tracking), GridMedia.cpp:71-98, on an OCS ray
// (o, d) inside [tNear, tFar]; returns true on a real collision at *t_out.
// sigma_bar = avg(extinction * densityMultiplier); inv_max = 1 / max density.
template <class R>
inline bool ref_delta_track(const Grid& g, float inv_max, float sigma_bar, V3f o, V3f d, float tNear, float tFar,
                            R& rng, float* t_out) {
  float t = tNear;
  while (true) {
    float r = rng.next();
    float dist = std::log(1 - r) * inv_max / sigma_bar;
    t -= dist;
    if (t >= tFar) { *t_out = tFar; return false; }
    float dens = ref_density_ocs(g, o + t * d);
    float ra = rng.next();
    if (dens * inv_max > ra) { *t_out = t; return true; }
  }
}
// GridMedia::Tr (ratio tracking with Russian roulette), GridMedia.cpp:45-69.
template <class R>
inline float ref_ratio_track(const Grid& g, float inv_max, float sigma_bar, V3f o, V3f d, float tNear, float tFar,
                             R& rng) {
  o = o + tNear * d;
  tFar = tFar - tNear;
  float Tr = 1, t = 0;
  while (true) {
    t -= std::log(1 - rng.next()) * inv_max / sigma_bar;
    if (t >= tFar) break;
    float dens = ref_density_ocs(g, o + t * d);
    Tr *= 1 - std::max(0.0f, dens * inv_max);
    const float rr = .1f;
    if (Tr < rr) {
      float q = std::max(0.05f, 1.0f - Tr);
      if (rng.next() < q) return 0.0f;
      Tr /= 1 - q;
    }
  }
  return Tr;
}

// HomogeneousMedia::Tr, src/materials/HomogeneousMedia.cpp:15-22:
// exp(-extinction * distance * density) per channel.
inline V3f ref_homogeneous_tr(V3f extinction, float density, float distance) {
  V3f v = extinction * distance * density;
  return V3f(std::exp(-v.x), std::exp(-v.y), std::exp(-v.z));
}
// HomogeneousMedia::sample, src/materials/HomogeneousMedia.cpp:24-50, as the
// reference writes it: the free-flight distance ignores the density and is
// divided by the chord length (tFar - tNear), and the weight is
// Tr * scattering / avg(extinction * Tr) on a collision, Tr / avg(Tr) else.
// Returns the collision flag; *t_out is the sampled t (or the chord length).
template <class R>
inline bool ref_homogeneous_sample(V3f scattering, V3f absorption, float density, float tNear, float tFar, R& rng,
                                   float* t_out, V3f* weight) {
  V3f ext = absorption + scattering;
  float t = -std::log(1 - rng.next()) / avg(ext);
  float dist = tFar - tNear;
  t = t / dist;
  bool sampled = t < dist;
  V3f tr = ref_homogeneous_tr(ext, density, t);
  V3f dens = sampled ? ext * tr : tr;
  float pdf = avg(dens);
  if (pdf == 0) pdf = 1;
  *weight = sampled ? tr * scattering / pdf : tr / pdf;
  *t_out = sampled ? t : dist;
  return sampled;
}

// Distribution1D, src/utils/Sampling.h:5-73: a piecewise-constant function of
// n values on [0, 1); cdf[i] = sum_{k<i} f[k] / n normalised by the integral
// (uniform when it is 0); sampleContinuous finds the last cdf entry <= u
// (binarySearch, src/utils/Math.h:1128-1145, clamped to [0, n - 1]).
struct RefDistribution1D {
  std::vector<float> func, cdf;
  float funcInt = 0;
  RefDistribution1D(const float* f, int n) : func(f, f + n), cdf(size_t(n) + 1) {
    cdf[0] = 0;
    for (int i = 1; i < n + 1; ++i) cdf[size_t(i)] = cdf[size_t(i) - 1] + func[size_t(i) - 1] / float(n);
    funcInt = cdf[size_t(n)];
    if (funcInt == 0)
      for (int i = 1; i < n + 1; ++i) cdf[size_t(i)] = float(i) / float(n);
    else
      for (int i = 1; i < n + 1; ++i) cdf[size_t(i)] /= funcInt;
  }
  int count() const { return int(func.size()); }
  float sampleContinuous(float u, float* pdf, int* off = nullptr) const {
    int first = 0, len = int(cdf.size());  // binarySearch(size, pred)
    while (len > 0) {
      int half = len >> 1, middle = first + half;
      if (cdf[size_t(middle)] <= u) { first = middle + 1; len -= half + 1; }
      else len = half;
    }
    int offset = std::min(std::max(first - 1, 0), int(cdf.size()) - 2);
    if (off) *off = offset;
    float du = u - cdf[size_t(offset)];
    if ((cdf[size_t(offset) + 1] - cdf[size_t(offset)]) > 0) du /= (cdf[size_t(offset) + 1] - cdf[size_t(offset)]);
    if (pdf) *pdf = (funcInt > 0) ? func[size_t(offset)] / funcInt : 0;
    return (float(offset) + du) / float(count());
  }
};
// Distribution2D, src/utils/Sampling.h:76-112.  The reference builds each
// row's conditional from `height` values (Sampling.h:87); cond_n selects
// that count (the reference's height, or width for the corrected table).
// With cond_n > width the reference reads past the row, so only
// cond_n <= width is restated.
struct RefDistribution2D {
  std::vector<RefDistribution1D> cond;
  std::unique_ptr<RefDistribution1D> marg;
  RefDistribution2D(const float* f, int width, int height, int cond_n) {
    cond.reserve(size_t(height));
    for (int v = 0; v < height; ++v) cond.emplace_back(f + size_t(v) * size_t(width), cond_n);
    std::vector<float> m;
    for (int v = 0; v < height; ++v) m.push_back(cond[size_t(v)].funcInt);
    marg.reset(new RefDistribution1D(m.data(), height));
  }
  void sampleContinuous(float u0, float u1, float* uv, float* pdf) const {
    float p1, p0;
    int v;
    uv[1] = marg->sampleContinuous(u1, &p1, &v);
    uv[0] = cond[size_t(v)].sampleContinuous(u0, &p0);
    *pdf = p0 * p1;
  }
};

// Sampling tables of an environment map (the corrected Distribution2D: each
// row's conditional over the row's own W texels): func = luminance *
// sin(theta of the row centre), FP64 sums, FP32 tables with exact 0 and 1
// ends, uniform where a row or the map is black; inv_int = W H / sum(func).
struct EnvTables {
  std::vector<float> func, cdf, marg;
  float inv_int = 0.0f;
  void build(const float* rgb, int W, int H) {
    func.assign(size_t(W) * H, 0.0f);
    cdf.assign(size_t(H) * (W + 1), 0.0f);
    marg.assign(size_t(H) + 1, 0.0f);
    std::vector<double> row(static_cast<size_t>(H));
    double total = 0.0;
    for (int j = 0; j < H; ++j) {
      const double st = std::sin(kPI * (double(j) + 0.5) / double(H));
      std::vector<double> pre(size_t(W) + 1, 0.0);
      double acc = 0.0;
      for (int i = 0; i < W; ++i) {
        const float* p = rgb + 3 * (size_t(j) * W + i);
        const float f = float((0.2126 * p[0] + 0.7152 * p[1] + 0.0722 * p[2]) * st);
        func[size_t(j) * W + i] = f;
        acc += double(f);
        pre[size_t(i) + 1] = acc;
      }
      float* c = &cdf[size_t(j) * (W + 1)];
      for (int i = 1; i < W; ++i) c[i] = acc > 0 ? float(pre[size_t(i)] / acc) : float(double(i) / W);
      c[0] = 0.0f;
      c[W] = 1.0f;
      row[size_t(j)] = acc;
      total += acc;
    }
    double acc = 0.0;
    for (int j = 1; j < H; ++j) {
      acc += row[size_t(j) - 1];
      marg[size_t(j)] = total > 0 ? float(acc / total) : float(double(j) / H);
    }
    marg[0] = 0.0f;
    marg[size_t(H)] = 1.0f;
    inv_int = total > 0 ? float(double(W) * double(H) / total) : 0.0f;
  }
};
inline int cdf_find(const float* cdf, int n, float u) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (cdf[mid] <= u) lo = mid;
    else hi = mid;
  }
  return lo;
}

// OfflineEngine::postProcessing, src/core/OfflineEngine.cpp:27-52.
inline V3f ref_post_processing(V3f c) {
  const float exposure = 0.5f;
  V3f m(1.0f - std::exp(-c.x * exposure), 1.0f - std::exp(-c.y * exposure), 1.0f - std::exp(-c.z * exposure));
  m = V3f(std::pow(m.x, 1.0f / 2.2f), std::pow(m.y, 1.0f / 2.2f), std::pow(m.z, 1.0f / 2.2f));
  return vmin(vmax(m, V3f(0.0f)), V3f(1.0f));
}

// Camera, src/core/Camera.cpp:7-26 and getRayPassingThrough :140-144 with the
// lens offset at zero (pinhole).
struct Camera {
  V3f position, lowerLeft, horizontal, vertical;
};
inline Camera ref_camera(V3f lookFrom, V3f lookAt, V3f up, float vfov_deg, float aspect, float focus) {
  Camera c;
  float halfHeight = std::tan(float(vfov_deg * kPI / 180.0) / 2.0f);
  float halfWidth = aspect * halfHeight;
  V3f front = normalize(lookFrom - lookAt);
  V3f side = normalize(cross(up, front));
  c.position = lookFrom;
  c.lowerLeft = lookFrom - focus * halfWidth * (-side) - focus * halfHeight * up - focus * (-front);
  c.horizontal = 2.0f * focus * halfWidth * (-side);
  c.vertical = 2.0f * focus * halfHeight * up;
  return c;
}
inline V3f ref_camera_dir(const Camera& c, float u, float v) {
  return -normalize(c.lowerLeft + u * c.horizontal + v * c.vertical - c.position);
}

// OfflineEngine tile geometry, src/core/OfflineEngine.cpp:18-21,54-58:
// 40 x 10 tiles of (W/40) x (H/10) pixels (integer division).
struct TileRect { int x0, y0, x1, y1; };
inline TileRect ref_tile_rect(int W, int H, int index) {
  int tw = W / 40, th = H / 10;
  int mx = (index % 40) * tw, my = (index / 40) * th;
  return {mx, my, mx + tw, my + th};
}

// ---------------------------------------------------------------------------
// 2. Scene description (flattened, world space) and the corrected integrator
// ---------------------------------------------------------------------------
// Material record: [type, albedo r g b, roughness, metallic, Le r g b, pad, pad]
enum { MAT_GLOSSY = 0, MAT_EMITTER = 1 };
enum { ENV_NONE = 0, ENV_CONSTANT = 1, ENV_SKY = 2, ENV_MAP = 3 };

struct SceneDesc {
  const float* tri_v; const int32_t* tri_mat; uint32_t n_tri;       // 9 floats per triangle
  const float* sph; const int32_t* sph_mat; uint32_t n_sph;         // cx cy cz r
  const float* rect; const int32_t* rect_mat; uint32_t n_rect;      // p0, eu, ev, n (12 floats)
  const float* mat; uint32_t n_mat;                                 // 12 floats per material
  int has_volume;
  float vol_xf[16], vol_inv[16];                                    // row-major OCS->WCS and inverse
  int vol_res[3];
  const float* grid;
  float vol_mult, vol_sigma_s[3], vol_sigma_a[3], vol_g;
  int maj_res;                                                      // majorant cells per axis (1 = global)
  int has_sun; float sun_dir[3]; float sun_E[3];                    // sun_dir: direction light travels
  int env_type; float env_c0[3], env_c1[3], env_c2[3];              // const: c0; sky: zenith c0, horizon c1, ground c2
  int vol_homogeneous;                                              // constant density vol_mult; grid unused
  const float* env_map; int env_w, env_h;                           // ENV_MAP: RGB, row 0 at theta = 0
  const float* point; const int32_t* point_mat; uint32_t n_point;   // point lights (emitter materials)
  int maj_no_skip;                                                  // device DDA option; no effect on results
};

struct CameraDesc {
  float position[3], look_at[3], up[3];
  float vfov, aspect;
};

template <class T>
struct Mat4 {
  T m[16];
  V3<T> point(V3<T> p) const {
    return V3<T>(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
                 m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
  }
  V3<T> vec(V3<T> p) const {
    return V3<T>(m[0] * p.x + m[1] * p.y + m[2] * p.z, m[4] * p.x + m[5] * p.y + m[6] * p.z,
                 m[8] * p.x + m[9] * p.y + m[10] * p.z);
  }
};

// Midpoint-split BVH over all surface primitives, following the reference
// builder's policy (src/primitives/BVH.cpp:6-106: split at the centre of the
// longest centroid axis, median fallback, <= 4 primitives per leaf).
struct BvhNode {
  float bmin[3], bmax[3];
  uint32_t start, count, right;  // leaf if right == 0 (children: i+1, i+right)
};
struct PrimRef { uint32_t type, index; };  // type 0 tri, 1 sphere, 2 rect

template <class T>
struct World {
  using V = V3<T>;
  const SceneDesc* s;
  std::vector<BvhNode> nodes;
  std::vector<PrimRef> prims;
  Grid grid;
  Mat4<T> xf, inv;
  int mres = 1;
  std::vector<float> maj;      // majorant of sigma_t per cell (mres^3)
  int occ_lo[3] = {0, 0, 0}, occ_hi[3] = {-1, -1, -1};  // cells with a positive majorant lie in [occ_lo, occ_hi]
  float sigma_bar = 0;         // avg(sigma_t * mult)
  EnvTables env;               // environment-map sampling tables
  int n_lights = 0;            // sun? + env? + emissive primitives
  struct Emitter { int type; uint32_t index; };  // type 0 tri, 1 sphere, 2 rect, 3 point; input index
  std::vector<Emitter> emitters;

  static V ldv(const float* q) { return V(T(q[0]), T(q[1]), T(q[2])); }

  explicit World(const SceneDesc* sd) : s(sd) {
    build_bvh();
    if (s->has_volume) {
      static const float kOne = 1.0f;
      if (s->vol_homogeneous) { grid.W = grid.H = grid.D = 1; grid.data = &kOne; }
      else { grid.W = s->vol_res[0]; grid.H = s->vol_res[1]; grid.D = s->vol_res[2]; grid.data = s->grid; }
      for (int i = 0; i < 16; ++i) { xf.m[i] = T(s->vol_xf[i]); inv.m[i] = T(s->vol_inv[i]); }
      V3f st(s->vol_sigma_a[0] + s->vol_sigma_s[0], s->vol_sigma_a[1] + s->vol_sigma_s[1],
             s->vol_sigma_a[2] + s->vol_sigma_s[2]);
      sigma_bar = avg(st * s->vol_mult);
      mres = s->vol_homogeneous ? 1 : std::max(1, s->maj_res);
      build_majorants();
    }
    if (s->env_type == ENV_MAP) env.build(s->env_map, s->env_w, s->env_h);
    if (s->has_sun) n_lights++;
    if (s->env_type != ENV_NONE) n_lights++;
    // light-selection order: emissive rectangles, spheres, triangles, each in input order
    auto emits = [&](int m) { return s->mat[12 * size_t(m)] == float(MAT_EMITTER); };
    for (uint32_t i = 0; i < s->n_rect; ++i)
      if (emits(s->rect_mat[i])) emitters.push_back({2, i});
    for (uint32_t i = 0; i < s->n_sph; ++i)
      if (emits(s->sph_mat[i])) emitters.push_back({1, i});
    for (uint32_t i = 0; i < s->n_tri; ++i)
      if (emits(s->tri_mat ? s->tri_mat[i] : 0)) emitters.push_back({0, i});
    for (uint32_t i = 0; i < s->n_point; ++i) emitters.push_back({3, i});
    n_lights += int(emitters.size());
  }

  // Majorant per cell: max voxel over the cell's trilinear support (the
  // cell's voxels plus a one-voxel apron on each side), times sigma_bar.
  void build_majorants() {
    maj.assign(size_t(mres) * mres * mres, 0.0f);
    int res[3] = {grid.W, grid.H, grid.D};
    for (int cz = 0; cz < mres; ++cz)
      for (int cy = 0; cy < mres; ++cy)
        for (int cx = 0; cx < mres; ++cx) {
          int c[3] = {cx, cy, cz}, lo[3], hi[3];
          for (int a = 0; a < 3; ++a) {
            lo[a] = int((int64_t(c[a]) * res[a]) / mres);
            hi[a] = std::min(res[a] - 1, int((int64_t(c[a] + 1) * res[a] + mres - 1) / mres) + 1);
            lo[a] = std::max(0, lo[a] - 1);
          }
          float m = 0;
          for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
              for (int x = lo[0]; x <= hi[0]; ++x) m = std::max(m, grid.at(x, y, z));
          maj[(size_t(cz) * mres + cy) * mres + cx] = m * sigma_bar;
        }
    for (int a = 0; a < 3; ++a) { occ_lo[a] = mres; occ_hi[a] = -1; }
    for (int cz = 0; cz < mres; ++cz)
      for (int cy = 0; cy < mres; ++cy)
        for (int cx = 0; cx < mres; ++cx)
          if (maj[(size_t(cz) * mres + cy) * mres + cx] > 0.0f) {
            const int c[3] = {cx, cy, cz};
            for (int a = 0; a < 3; ++a) {
              occ_lo[a] = std::min(occ_lo[a], c[a]);
              occ_hi[a] = std::max(occ_hi[a], c[a]);
            }
          }
  }

  void prim_bounds(const PrimRef& p, float mn[3], float mx[3]) const {
    if (p.type == 0) {
      const float* v = s->tri_v + 9 * size_t(p.index);
      for (int a = 0; a < 3; ++a) {
        mn[a] = std::min(v[a], std::min(v[3 + a], v[6 + a]));
        mx[a] = std::max(v[a], std::max(v[3 + a], v[6 + a]));
      }
    } else if (p.type == 1) {
      const float* q = s->sph + 4 * size_t(p.index);
      for (int a = 0; a < 3; ++a) { mn[a] = q[a] - q[3]; mx[a] = q[a] + q[3]; }
    } else {
      const float* r = s->rect + 12 * size_t(p.index);
      for (int a = 0; a < 3; ++a) {
        float c0 = r[a], c1 = r[a] + r[3 + a], c2 = r[a] + r[6 + a], c3 = r[a] + r[3 + a] + r[6 + a];
        mn[a] = std::min(std::min(c0, c1), std::min(c2, c3));
        mx[a] = std::max(std::max(c0, c1), std::max(c2, c3));
      }
    }
  }

  void build_bvh() {
    for (uint32_t i = 0; i < s->n_tri; ++i) prims.push_back({0, i});
    for (uint32_t i = 0; i < s->n_sph; ++i) prims.push_back({1, i});
    for (uint32_t i = 0; i < s->n_rect; ++i) prims.push_back({2, i});
    if (prims.empty()) return;
    std::vector<float> cen(prims.size() * 3), pbmin(prims.size() * 3), pbmax(prims.size() * 3);
    for (size_t i = 0; i < prims.size(); ++i) {
      prim_bounds(prims[i], &pbmin[3 * i], &pbmax[3 * i]);
      for (int a = 0; a < 3; ++a) cen[3 * i + a] = 0.5f * (pbmin[3 * i + a] + pbmax[3 * i + a]);
    }
    std::vector<uint32_t> idx(prims.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = uint32_t(i);
    nodes.reserve(prims.size() * 2);
    struct Item { uint32_t parent, start, end; };
    std::vector<Item> st;
    st.push_back({0xffffffffu, 0, uint32_t(prims.size())});
    while (!st.empty()) {
      Item it = st.back(); st.pop_back();
      uint32_t ni = uint32_t(nodes.size());
      BvhNode n{};
      float bmn[3] = {kINF, kINF, kINF}, bmx[3] = {-kINF, -kINF, -kINF};
      float cmn[3] = {kINF, kINF, kINF}, cmx[3] = {-kINF, -kINF, -kINF};
      for (uint32_t i = it.start; i < it.end; ++i) {
        uint32_t p = idx[i];
        for (int a = 0; a < 3; ++a) {
          bmn[a] = std::min(bmn[a], pbmin[3 * p + a]); bmx[a] = std::max(bmx[a], pbmax[3 * p + a]);
          cmn[a] = std::min(cmn[a], cen[3 * p + a]); cmx[a] = std::max(cmx[a], cen[3 * p + a]);
        }
      }
      for (int a = 0; a < 3; ++a) { n.bmin[a] = bmn[a]; n.bmax[a] = bmx[a]; }
      n.start = it.start; n.count = it.end - it.start; n.right = 0;
      nodes.push_back(n);
      if (n.count <= 4) continue;
      int axis = 0;
      float ext[3] = {cmx[0] - cmn[0], cmx[1] - cmn[1], cmx[2] - cmn[2]};
      if (ext[1] > ext[axis]) axis = 1;
      if (ext[2] > ext[axis]) axis = 2;
      float split = 0.5f * (cmn[axis] + cmx[axis]);
      uint32_t mid = it.start;
      for (uint32_t i = it.start; i < it.end; ++i)
        if (cen[3 * idx[i] + axis] < split) { std::swap(idx[i], idx[mid]); ++mid; }
      if (mid == it.start || mid == it.end) mid = it.start + (it.end - it.start) / 2;
      // depth-first layout: left child = ni + 1; right child index patched when built
      nodes[ni].right = 0xffffffffu;
      st.push_back({ni, mid, it.end});   // right (built second)
      st.push_back({ni, it.start, mid}); // left (built first, lands at ni + 1)
    }
    // patch right-child offsets: a node's right child is the first node after
    // its left subtree.  Recompute by a second pass with explicit subtree sizes.
    std::vector<uint32_t> sub(nodes.size(), 1);
    for (size_t k = nodes.size(); k-- > 0;) {
      if (nodes[k].right == 0) continue;
      uint32_t l = uint32_t(k) + 1;
      uint32_t r = l + sub[l];
      nodes[k].right = r - uint32_t(k);
      sub[k] = 1 + sub[l] + sub[r];
    }
    std::vector<PrimRef> sorted(prims.size());
    for (size_t i = 0; i < idx.size(); ++i) sorted[i] = prims[idx[i]];
    prims.swap(sorted);
  }

  // --- surface intersection (corrected: interval (tmin, tmax), nearest root)
  struct SHit { T t; int type; uint32_t index; V n; int mat; };

  static bool slab(const BvhNode& b, V o, V invd, T tmax) {
    T t0 = T(0), t1 = tmax;
    for (int a = 0; a < 3; ++a) {
      T tn = (T(b.bmin[a]) - o[a]) * invd[a];
      T tf = (T(b.bmax[a]) - o[a]) * invd[a];
      if (tn > tf) std::swap(tn, tf);
      tf = tf * T(1.0000003576);  // 1 + 2 gamma(3) (robust slab, Ize 2013)
      t0 = tn > t0 ? tn : t0;
      t1 = tf < t1 ? tf : t1;
      if (t0 > t1) return false;
    }
    return true;
  }

  bool isect_prim(const PrimRef& p, V o, V d, T tmin, T tmax, T& t_out) const {
    if (p.type == 0) {
      const float* q = s->tri_v + 9 * size_t(p.index);
      V v0 = ldv(q), v1 = ldv(q + 3), v2 = ldv(q + 6);
      V e1 = v1 - v0, e2 = v2 - v0, rov0 = o - v0;
      V n = cross(e1, e2), qq = cross(rov0, d);
      T inv = T(1) / dot(d, n);
      T u = dot(-qq, e2) * inv, v = dot(qq, e1) * inv, t = dot(-n, rov0) * inv;
      if (!(t > tmin && t < tmax)) return false;
      if (u < T(0) || u > T(1) || v < T(0) || (u + v) > T(1)) return false;
      t_out = t;
      return true;
    } else if (p.type == 1) {
      const float* q = s->sph + 4 * size_t(p.index);
      V c = ldv(q);
      T r = T(q[3]);
      V oc = o - c;
      T a = dot(d, d), b = dot(oc, d), cc = dot(oc, oc) - r * r;
      T disc = b * b - a * cc;
      if (disc < T(0)) return false;
      T sq = std::sqrt(disc);
      T t = (-b - sq) / a;
      if (!(t > tmin)) t = (-b + sq) / a;
      if (!(t > tmin && t < tmax)) return false;
      t_out = t;
      return true;
    } else {
      const float* q = s->rect + 12 * size_t(p.index);
      V p0 = ldv(q), eu = ldv(q + 3), ev = ldv(q + 6);
      V n = ldv(q + 9);
      T den = dot(n, d);
      if (std::fabs(den) < T(kEPS)) return false;
      T t = dot(n, p0 - o) / den;
      if (!(t > tmin && t < tmax)) return false;
      V h = (o + t * d) - p0;
      T a = dot(h, eu) / dot(eu, eu), b = dot(h, ev) / dot(ev, ev);
      if (a < T(0) || a > T(1) || b < T(0) || b > T(1)) return false;
      t_out = t;
      return true;
    }
  }

  // closest hit; ties on t resolve to the lower (type, index) key
  bool closest(V o, V d, T tmin, T tmax, SHit& out) const {
    if (nodes.empty()) return false;
    V invd(T(1) / d.x, T(1) / d.y, T(1) / d.z);
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    bool found = false;
    T best = tmax;
    uint64_t best_key = ~0ull;
    while (sp > 0) {
      const BvhNode& nd = nodes[stack[--sp]];
      if (!slab(nd, o, invd, best)) continue;
      if (nd.right == 0) {
        for (uint32_t i = nd.start; i < nd.start + nd.count; ++i) {
          T t;
          if (isect_prim(prims[i], o, d, tmin, tmax, t)) {
            uint64_t key = (uint64_t(prims[i].type) << 32) | prims[i].index;
            if (t < best || (found && t == best && key < best_key)) {
              best = t; best_key = key; found = true;
              out.t = t; out.type = int(prims[i].type); out.index = prims[i].index;
            }
          }
        }
      } else {
        uint32_t ni = uint32_t(&nd - nodes.data());
        stack[sp++] = ni + nd.right;
        stack[sp++] = ni + 1;
      }
    }
    if (!found) return false;
    fill_normal(out, o, d);
    return true;
  }

  bool occluded(V o, V d, T tmin, T tmax) const {
    if (nodes.empty()) return false;
    V invd(T(1) / d.x, T(1) / d.y, T(1) / d.z);
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const BvhNode& nd = nodes[stack[--sp]];
      if (!slab(nd, o, invd, tmax)) continue;
      if (nd.right == 0) {
        for (uint32_t i = nd.start; i < nd.start + nd.count; ++i) {
          T t;
          if (isect_prim(prims[i], o, d, tmin, tmax, t)) return true;
        }
      } else {
        uint32_t ni = uint32_t(&nd - nodes.data());
        stack[sp++] = ni + nd.right;
        stack[sp++] = ni + 1;
      }
    }
    return false;
  }

  void fill_normal(SHit& h, V o, V d) const {
    if (h.type == 0) {
      const float* q = s->tri_v + 9 * size_t(h.index);
      V v0 = ldv(q), v1 = ldv(q + 3), v2 = ldv(q + 6);
      h.n = normalize(cross(v1 - v0, v2 - v0));
      h.mat = s->tri_mat ? s->tri_mat[h.index] : 0;
    } else if (h.type == 1) {
      const float* q = s->sph + 4 * size_t(h.index);
      V c = ldv(q);
      h.n = ((o + h.t * d) - c) / T(q[3]);
      h.mat = s->sph_mat[h.index];
    } else {
      const float* q = s->rect + 12 * size_t(h.index);
      h.n = V(T(q[9]), T(q[10]), T(q[11]));
      h.mat = s->rect_mat[h.index];
    }
  }

  // --- volume: OCS box interval
  bool box_interval(V o, V d, T& ta, T& tb) const {
    V oo = inv.point(o), dd = inv.vec(d);
    T t0 = -std::numeric_limits<T>::infinity(), t1 = std::numeric_limits<T>::infinity();
    for (int a = 0; a < 3; ++a) {
      T id = T(1) / dd[a];
      T tn = (T(-0.5) - oo[a]) * id, tf = (T(0.5) - oo[a]) * id;
      if (tn > tf) std::swap(tn, tf);
      t0 = tn > t0 ? tn : t0;
      t1 = tf < t1 ? tf : t1;
    }
    ta = t0; tb = t1;
    return t0 <= t1;
  }

  // A homogeneous medium has density 1 (times vol_mult) everywhere in the cube.
  T density_ocs(V p) const {
    if (s->vol_homogeneous) return T(1);
    V3f pf(float(p.x), float(p.y), float(p.z));
    return T(ref_density_ocs(grid, pf));
  }

  // Visit majorant cells along the OCS ray over [ta, tb]; f(t0, t1, mu)
  // returns false to stop.
  template <class F>
  void dda(V oo, V dd, T ta, T tb, F&& f) const {
    const T M = T(mres), invM = T(1) / M;
    int cell[3], step[3];
    T tnext[3], inv[3];
    // The walk is clipped to the box of the cells with a positive majorant:
    // its faces are cell boundaries in the walk's own arithmetic, so the cells
    // and boundary distances inside are those of a walk from the volume box.
    for (int a = 0; a < 3; ++a) {
      inv[a] = dd[a] != T(0) ? T(1) / dd[a] : T(0);  // cell boundaries: multiply by 1 / d
      const T blo = T(occ_lo[a]) * invM - T(0.5), bhi = T(occ_hi[a] + 1) * invM - T(0.5);
      if (occ_lo[a] > occ_hi[a]) {
        tb = ta;
      } else if (dd[a] != T(0)) {
        const T t0 = (blo - oo[a]) * inv[a], t1 = (bhi - oo[a]) * inv[a];
        ta = std::max(ta, std::min(t0, t1));
        tb = std::min(tb, std::max(t0, t1));
      } else if (!(oo[a] >= blo && oo[a] <= bhi)) {
        tb = ta;
      }
    }
    if (!(ta < tb)) return;
    V pe = oo + ta * dd;
    for (int a = 0; a < 3; ++a) {
      T g = (pe[a] + T(0.5)) * M;
      int c = int(std::floor(g));
      c = std::min(std::max(c, 0), mres - 1);
      cell[a] = c;
      if (dd[a] > T(0)) { step[a] = 1; tnext[a] = ((T(c + 1) * invM - T(0.5)) - oo[a]) * inv[a]; }
      else if (dd[a] < T(0)) { step[a] = -1; tnext[a] = ((T(c) * invM - T(0.5)) - oo[a]) * inv[a]; }
      else { step[a] = 0; tnext[a] = std::numeric_limits<T>::infinity(); }
    }
    T t = ta;
    while (t < tb) {
      int ax = 0;
      if (tnext[1] < tnext[ax]) ax = 1;
      if (tnext[2] < tnext[ax]) ax = 2;
      T tc = std::min(tnext[ax], tb);
      T mu = T(maj[(size_t(cell[2]) * mres + cell[1]) * mres + cell[0]]);
      if (tc > t && !f(t, tc, mu)) return;
      t = tc;
      if (tnext[ax] >= tb) break;
      cell[ax] += step[ax];
      if (cell[ax] < 0 || cell[ax] >= mres) break;
      T bnd = step[ax] > 0 ? T(cell[ax] + 1) * invM : T(cell[ax]) * invM;
      tnext[ax] = ((bnd - T(0.5)) - oo[ax]) * inv[ax];
    }
  }

  // Delta tracking over [ta, tb] with per-cell majorants (a 1^3 grid is the
  // reference's global majorant).  Same step law and acceptance test as
  // GridMedia::sample (GridMedia.cpp:80-96), which multiplies by the inverse
  // majorant: t -= log(1 - xi) * (1 / mu), accept if sigma * (1 / mu) > xi'.
  bool delta_track(V o, V d, T ta, T tb, Rng& rng, T& t_hit) const {
    V oo = inv.point(o), dd = inv.vec(d);
    bool hit = false;
    dda(oo, dd, ta, tb, [&](T c0, T c1, T mu) {
      if (mu <= T(0)) return true;
      const T imu = T(1) / mu;
      T t = c0;
      while (true) {
        T r = T(rng.next());
        t -= std::log(T(1) - r) * imu;
        if (t >= c1) return true;
        T sig = density_ocs(oo + t * dd) * T(sigma_bar);
        T ra = T(rng.next());
        if (sig * imu > ra) { hit = true; t_hit = t; return false; }
      }
    });
    return hit;
  }

  // Ratio tracking with the reference's Russian roulette (GridMedia.cpp:54-66).
  // In a homogeneous medium the ratio-tracking expectation is taken in closed
  // form: exp(-sigma_bar (tb - ta)), HomogeneousMedia::Tr with the grey
  // extinction of the tracking step law.
  T ratio_track(V o, V d, T ta, T tb, Rng& rng) const {
    if (s->vol_homogeneous) return std::exp(-T(sigma_bar) * (tb - ta));
    V oo = inv.point(o), dd = inv.vec(d);
    T Tr = T(1);
    dda(oo, dd, ta, tb, [&](T c0, T c1, T mu) {
      if (mu <= T(0)) return true;
      const T imu = T(1) / mu;
      T t = c0;
      while (true) {
        t -= std::log(T(1) - T(rng.next())) * imu;
        if (t >= c1) return true;
        T sig = density_ocs(oo + t * dd) * T(sigma_bar);
        Tr *= T(1) - std::max(T(0), sig * imu);
        if (Tr < T(0.1)) {
          T q = std::max(T(0.05), T(1) - Tr);
          if (T(rng.next()) < q) { Tr = T(0); return false; }
          Tr /= T(1) - q;
        }
      }
    });
    return Tr;
  }

  // --- lights
  // Environment map lookup as InfiniteAreaLight::Le (InfiniteAreaLight.h:47-55):
  // local frame of (0,0,1) from generateOrthonormalCS, (x, y, z) = (d.y, -d.x, d.z),
  // u = phi / 2 pi (phi in [0, 2 pi)), v = theta / pi; piecewise-constant texels.
  int env_texel(V d, T& sin_t) const {
    T phi = std::atan2(-d.x, d.y);
    phi = phi < T(0) ? phi + T(kTWO_PI) : phi;
    T z = std::min(T(1), std::max(T(-1), d.z));
    sin_t = std::sqrt(std::max(T(0), T(1) - z * z));
    T theta = std::acos(z);
    int i = std::min(int(phi * T(1.0 / kTWO_PI) * T(s->env_w)), s->env_w - 1);
    int j = std::min(int(theta * T(1.0 / kPI) * T(s->env_h)), s->env_h - 1);
    return j * s->env_w + i;
  }
  T env_pdf(V d) const {
    if (s->env_type != ENV_MAP) return T(1.0 / kFOUR_PI);
    T st;
    int k = env_texel(d, st);
    if (!(st > T(0))) return T(0);
    return T(env.func[size_t(k)]) * T(env.inv_int) / (T(2.0 * kPI * kPI) * st);
  }
  void env_sample(T ua, T ub, V& wi, T& pl, V& Le) const {
    if (s->env_type != ENV_MAP) {
      T z = T(1) - T(2) * ua, r = std::sqrt(std::max(T(0), T(1) - z * z)), ph = T(kTWO_PI) * ub;
      wi = V(r * std::cos(ph), r * std::sin(ph), z);
      Le = env_le(wi);
      pl = T(1.0 / kFOUR_PI);
      return;
    }
    const int W = s->env_w, H = s->env_h;
    const float fv = float(ub), fu = float(ua);
    int j = cdf_find(env.marg.data(), H, fv);
    float m0 = env.marg[size_t(j)], m1 = env.marg[size_t(j) + 1];
    float dv = m1 > m0 ? (fv - m0) / (m1 - m0) : 0.0f;
    const float* row = env.cdf.data() + size_t(j) * size_t(W + 1);
    int i = cdf_find(row, W, fu);
    float c0 = row[i], c1 = row[i + 1];
    float du = c1 > c0 ? (fu - c0) / (c1 - c0) : 0.0f;
    T u = (T(i) + T(du)) / T(W), v = (T(j) + T(dv)) / T(H);
    T theta = v * T(kPI), phi = u * T(kTWO_PI);
    T st = std::sin(theta), ct = std::cos(theta);
    wi = V(-st * std::sin(phi), st * std::cos(phi), ct);
    size_t k = size_t(j) * size_t(W) + size_t(i);
    pl = st > T(0) ? T(env.func[k]) * T(env.inv_int) / (T(2.0 * kPI * kPI) * st) : T(0);
    Le = ldv(s->env_map + 3 * k);
  }
  V env_le(V d) const {
    if (s->env_type == ENV_MAP) {
      T st;
      return ldv(s->env_map + 3 * size_t(env_texel(d, st)));
    }
    if (s->env_type == ENV_CONSTANT) return V(T(s->env_c0[0]), T(s->env_c0[1]), T(s->env_c0[2]));
    if (s->env_type == ENV_SKY) {
      V zen(T(s->env_c0[0]), T(s->env_c0[1]), T(s->env_c0[2]));
      V hor(T(s->env_c1[0]), T(s->env_c1[1]), T(s->env_c1[2]));
      V gnd(T(s->env_c2[0]), T(s->env_c2[1]), T(s->env_c2[2]));
      if (d.y >= T(0)) return hor * (T(1) - d.y) + zen * d.y;
      return gnd;
    }
    return V(T(0));
  }

  // Emissive primitives (corrected forms; the reference's sphere sampler
  // returns the nearest surface point with a mis-scaled area pdf,
  // src/primitives/Sphere.cpp:46-67):
  //   rectangle: uniform point, area pdf d^2 / (cos area)
  //   triangle:  uniform point (b1 = v sqrt(u), b2 = sqrt(u)(1 - v)), area pdf
  //   sphere:    uniform direction in the subtended cone, pdf 1 / (2 pi (1 - cos theta_max));
  //              uniform surface point when x is inside.
  static T cone_width(T dc2, T r2) {
    T sin2 = r2 / dc2;
    return sin2 / (T(1) + std::sqrt(std::max(T(0), T(1) - sin2)));
  }
  // Point lights: delta lights of intensity = the emitter's radiance, arriving
  // as intensity / d^2 (the reference's Point pdf is 1 with no falloff,
  // src/primitives/Point.cpp:14-26).  Returns true for a delta light.
  bool sample_emitter(const Emitter& e, V x, T ua, T ub, V& wi, T& dist, T& pl, V& Le) const {
    int m;
    if (e.type == 3) {
      V dv = ldv(s->point + 3 * size_t(e.index)) - x;
      T d2 = dot(dv, dv);
      dist = std::sqrt(d2);
      wi = dv / dist;
      pl = d2 > T(0) ? T(1) : T(0);
      Le = ldv(s->mat + 12 * size_t(s->point_mat[e.index]) + 6) / d2;
      return true;
    }
    if (e.type == 2) {
      const float* q = s->rect + 12 * size_t(e.index);
      V p0 = ldv(q), eu = ldv(q + 3), ev = ldv(q + 6), ln = ldv(q + 9);
      V y = p0 + ua * eu + ub * ev;
      V dv = y - x;
      T d2 = dot(dv, dv);
      dist = std::sqrt(d2);
      wi = dv / dist;
      T cosl = std::fabs(dot(ln, wi));
      T area = length(cross(eu, ev));
      m = s->rect_mat[e.index];
      pl = cosl > T(0) ? d2 / (cosl * area) : T(0);
    } else if (e.type == 1) {
      const float* q = s->sph + 4 * size_t(e.index);
      V c = ldv(q);
      T r = T(q[3]), r2 = r * r;
      V oc = c - x;
      T dc2 = dot(oc, oc);
      m = s->sph_mat[e.index];
      if (dc2 > r2) {
        T dc = std::sqrt(dc2);
        V w = oc / dc;
        T wid = cone_width(dc2, r2);
        T qq = ua * wid;
        T ct = T(1) - qq, st = std::sqrt(std::max(T(0), qq * (T(2) - qq)));
        T ph = T(kTWO_PI) * ub;
        V u, v;
        ref_onb(w, u, v);
        wi = u * (st * std::cos(ph)) + v * (st * std::sin(ph)) + w * ct;
        dist = dc * ct - std::sqrt(std::max(T(0), r2 - dc2 * st * st));
        pl = T(1) / (T(kTWO_PI) * wid);
      } else {
        T z = T(1) - T(2) * ua, rr = std::sqrt(std::max(T(0), T(1) - z * z)), ph = T(kTWO_PI) * ub;
        V nl(rr * std::cos(ph), rr * std::sin(ph), z);
        V dv = (c + r * nl) - x;
        T d2 = dot(dv, dv);
        dist = std::sqrt(d2);
        wi = dv / dist;
        T cosl = std::fabs(dot(nl, wi));
        pl = cosl > T(0) ? d2 / (cosl * T(2.0 * kTWO_PI) * r2) : T(0);
      }
    } else {
      const float* q = s->tri_v + 9 * size_t(e.index);
      V v0 = ldv(q), e1 = ldv(q + 3) - v0, e2 = ldv(q + 6) - v0;
      T su = std::sqrt(ua);
      V y = v0 + (ub * su) * e1 + (su * (T(1) - ub)) * e2;
      V dv = y - x;
      T d2 = dot(dv, dv);
      dist = std::sqrt(d2);
      wi = dv / dist;
      V cr = cross(e1, e2);
      T len = length(cr);
      T cosl = std::fabs(dot(cr, wi)) / len;
      m = s->tri_mat ? s->tri_mat[e.index] : 0;
      pl = cosl > T(0) ? d2 / (cosl * (T(0.5) * len)) : T(0);
    }
    Le = ldv(s->mat + 12 * size_t(m) + 6);
    return false;
  }
  // Solid-angle pdf of sample_emitter for the point the ray (o, d) hits.
  T emitter_hit_pdf(const SHit& h, V o, V d) const {
    T cosl = std::fabs(dot(h.n, d));
    if (h.type == 2) {
      const float* q = s->rect + 12 * size_t(h.index);
      T area = length(cross(ldv(q + 3), ldv(q + 6)));
      return (h.t * h.t) / (cosl * area);
    }
    if (h.type == 1) {
      const float* q = s->sph + 4 * size_t(h.index);
      T r2 = T(q[3]) * T(q[3]);
      V oc = ldv(q) - o;
      T dc2 = dot(oc, oc);
      if (dc2 > r2) return T(1) / (T(kTWO_PI) * cone_width(dc2, r2));
      return (h.t * h.t) / (cosl * T(2.0 * kTWO_PI) * r2);
    }
    const float* q = s->tri_v + 9 * size_t(h.index);
    V v0 = ldv(q);
    T len = length(cross(ldv(q + 3) - v0, ldv(q + 6) - v0));
    return (h.t * h.t) / (cosl * (T(0.5) * len));
  }

  T transmittance(V o, V d, T tmax, Rng& rng) const {
    if (occluded(o, d, T(1e-4), tmax)) return T(0);
    if (!s->has_volume) return T(1);
    T ta, tb;
    if (!box_interval(o, d, ta, tb)) return T(1);
    ta = std::max(ta, T(0));
    tb = std::min(tb, tmax);
    if (ta >= tb) return T(1);
    return ratio_track(o, d, ta, tb, rng);
  }
};

// --- materials (corrected GGX + Lambert, two-sided) ------------------------
template <class T>
struct Glossy {
  V3<T> albedo; T alpha, metallic;
  static T D(T alpha, T cos_h) {
    T a2 = alpha * alpha, c2 = cos_h * cos_h;
    T den = c2 * (a2 - T(1)) + T(1);
    den = std::max(T(kEPS), T(kPI) * den * den);
    return a2 / den;
  }
  static T G1(T alpha, T c) {
    T k = alpha / T(2);
    return c / std::max(T(kEPS), c * (T(1) - k) + k);
  }
  static T F(T c) { T m = T(1) - c; T m2 = m * m; return T(0.04) + T(0.96) * (m2 * m2 * m); }
  T p_spec(T cos_o) const {
    T fs = F(cos_o);
    T kd = avg(albedo) * (T(1) - fs) * (T(1) - metallic);
    T p = fs / (fs + kd);
    return std::min(T(0.9), std::max(T(0.1), p));
  }
  // f (without the cosine), local frame z = n
  V3<T> eval(V3<T> wo, V3<T> wi) const {
    if (wo.z <= T(0) || wi.z <= T(0)) return V3<T>(T(0));
    V3<T> h = normalize(wo + wi);
    T Fh = F(dot(wo, h));
    T spec = D(alpha, h.z) * G1(alpha, wo.z) * G1(alpha, wi.z) * Fh / (T(4) * wo.z * wi.z);
    T kd = (T(1) - Fh) * (T(1) - metallic);
    return albedo * (kd / T(kPI)) + V3<T>(spec);
  }
  T pdf(V3<T> wo, V3<T> wi) const {
    if (wo.z <= T(0) || wi.z <= T(0)) return T(0);
    V3<T> h = normalize(wo + wi);
    T ps = p_spec(wo.z);
    T pspec = D(alpha, h.z) * h.z / (T(4) * std::fabs(dot(wo, h)));
    T pdiff = wi.z / T(kPI);
    return ps * pspec + (T(1) - ps) * pdiff;
  }
  // sample: u0 chooses the lobe, (u1, u2) the direction
  V3<T> sample(V3<T> wo, T u0, T u1, T u2) const {
    T ps = p_spec(wo.z);
    if (u0 < ps) {
      T a2 = alpha * alpha;
      T ct = std::min(T(1), std::sqrt((T(1) - u1) / std::max(T(kEPS), u1 * (a2 - T(1)) + T(1))));
      T st = std::sqrt(std::max(T(0), T(1) - ct * ct));
      T ph = T(kTWO_PI) * u2;
      V3<T> h(st * std::cos(ph), st * std::sin(ph), ct);
      return T(2) * dot(wo, h) * h - wo;
    }
    T r = std::sqrt(u1), ph = T(kTWO_PI) * u2;
    return V3<T>(r * std::cos(ph), r * std::sin(ph), std::sqrt(std::max(T(0), T(1) - u1)));
  }
};

template <class T>
inline T hg_p(T g, T cos_t) {
  T den = T(1) + g * g - T(2) * g * cos_t;
  return T(kINV4PI) * (T(1) - g * g) / (den * std::sqrt(den));
}
// forward-convention HG sample around propagation direction d
template <class T>
inline V3<T> hg_sample(T g, V3<T> d, T u0, T u1) {
  T ct;
  if (std::fabs(g) < T(1e-3)) ct = T(1) - T(2) * u0;
  else {
    T s = (T(1) - g * g) / (T(1) + g - T(2) * g * u0);
    ct = (T(1) + g * g - s * s) / (T(2) * g);
  }
  ct = std::min(T(1), std::max(T(-1), ct));
  T st = std::sqrt(std::max(T(0), T(1) - ct * ct));
  T ph = T(kTWO_PI) * u1;
  V3<T> ss, ts;
  ref_onb(d, ss, ts);
  return to_world(V3<T>(st * std::cos(ph), st * std::sin(ph), ct), d, ss, ts);
}

struct PathStats { uint64_t ext_rays = 0, shadow_rays = 0; };

// Corrected path integrator.  Event/draw order (mirrored by the GPU kernel):
//   camera jitter (2); per segment: delta-tracking draws; at a vertex:
//   light pick (1), light sample (2), shadow ratio-tracking draws,
//   direction sample (phase: 2; glossy: 3).
template <class T>
V3<T> Li(const World<T>& w, V3<T> o, V3<T> d, int max_bounces, Rng& rng, PathStats* st = nullptr) {
  using V = V3<T>;
  const SceneDesc* s = w.s;
  V L(T(0)), beta(T(1));
  T prev_pdf = T(0);
  bool prev_delta = true;
  int depth = 0;
  const T tmin = T(1e-4);
  const T g = T(s->vol_g);
  V sig_s(T(s->vol_sigma_s[0]), T(s->vol_sigma_s[1]), T(s->vol_sigma_s[2]));
  V sig_t(T(s->vol_sigma_s[0] + s->vol_sigma_a[0]), T(s->vol_sigma_s[1] + s->vol_sigma_a[1]),
          T(s->vol_sigma_s[2] + s->vol_sigma_a[2]));
  const T nl = T(w.n_lights);
  while (true) {
    if (st) st->ext_rays++;
    typename World<T>::SHit hit;
    bool hs = w.closest(o, d, tmin, std::numeric_limits<T>::infinity(), hit);
    T tS = hs ? hit.t : std::numeric_limits<T>::infinity();
    bool medium_event = false;
    V x, n;
    if (s->has_volume) {
      T ta, tb;
      if (w.box_interval(o, d, ta, tb)) {
        ta = std::max(ta, T(0));
        tb = std::min(tb, tS);
        T th;
        if (ta < tb && w.delta_track(o, d, ta, tb, rng, th)) { medium_event = true; x = o + th * d; }
      }
    }
    if (!medium_event) {
      if (!hs) {
        if (s->env_type != ENV_NONE) {
          T wgt = prev_delta ? T(1) : power_heuristic(prev_pdf, w.env_pdf(d));
          L = L + beta * w.env_le(d) * wgt;
        }
        break;
      }
      const float* m = s->mat + 12 * size_t(hit.mat);
      if (m[0] == float(MAT_EMITTER)) {
        V Le = World<T>::ldv(m + 6);
        T wgt = T(1);
        if (!prev_delta) wgt = power_heuristic(prev_pdf, w.emitter_hit_pdf(hit, o, d));
        L = L + beta * Le * wgt;
        break;
      }
      x = o + hit.t * d;
      n = hit.n;
    }
    if (depth >= max_bounces) break;
    // ---- vertex: NEE
    V wo = -d;
    Glossy<T> mat{};
    V ns, ss, ts;
    if (medium_event) {
      beta = beta * (sig_s / sig_t);
    } else {
      const float* m = s->mat + 12 * size_t(hit.mat);
      mat.albedo = World<T>::ldv(m + 1);
      T rough = T(m[4]);
      mat.alpha = rough * rough;
      mat.metallic = T(m[5]);
      ns = dot(n, wo) < T(0) ? -n : n;
      ref_onb(ns, ss, ts);
    }
    T ul = T(rng.next()), ua = T(rng.next()), ub = T(rng.next());
    if (w.n_lights > 0) {
      int li = std::min(int(ul * nl), w.n_lights - 1);
      int k = 0;
      V wi, Le(T(0));
      T pl = T(0), dist = std::numeric_limits<T>::infinity();
      bool delta = false;
      if (s->has_sun && li == k++) {
        wi = -V(T(s->sun_dir[0]), T(s->sun_dir[1]), T(s->sun_dir[2]));
        Le = V(T(s->sun_E[0]), T(s->sun_E[1]), T(s->sun_E[2]));
        pl = T(1); delta = true;
      } else if (s->env_type != ENV_NONE && li == (s->has_sun ? 1 : 0)) {
        w.env_sample(ua, ub, wi, pl, Le);
      } else {
        const auto& e = w.emitters[size_t(li - (s->has_sun ? 1 : 0) - (s->env_type != ENV_NONE ? 1 : 0))];
        delta = w.sample_emitter(e, x, ua, ub, wi, dist, pl, Le);
      }
      if (pl > T(0)) {
        V f; T pf;
        if (medium_event) {
          T p = hg_p(g, dot(d, wi));
          f = V(p); pf = p;
        } else {
          V wol = to_lcs(wo, ns, ss, ts), wil = to_lcs(wi, ns, ss, ts);
          f = mat.eval(wol, wil) * std::fabs(wil.z);
          pf = mat.pdf(wol, wil);
        }
        if (!is_black(f)) {
          if (st) st->shadow_rays++;
          T tmax = std::isinf(dist) ? std::numeric_limits<T>::infinity() : dist * T(1 - 1e-4);
          Rng srng = rng.shadow_stream(depth);
          T tr = w.transmittance(x, wi, tmax, srng);
          if (tr > T(0)) {
            T wgt = delta ? T(1) : power_heuristic(pl, pf);
            L = L + beta * f * Le * (tr * wgt * nl / pl);
          }
        }
      }
    }
    // ---- continue: sample direction
    if (medium_event) {
      T u0 = T(rng.next()), u1 = T(rng.next());
      V wi = hg_sample(g, d, u0, u1);
      prev_pdf = hg_p(g, dot(d, wi));
      d = wi;
    } else {
      T u0 = T(rng.next()), u1 = T(rng.next()), u2 = T(rng.next());
      V wol = to_lcs(wo, ns, ss, ts);
      V wil = mat.sample(wol, u0, u1, u2);
      T pf = mat.pdf(wol, wil);
      if (!(pf > T(0))) break;
      V f = mat.eval(wol, wil) * std::fabs(wil.z);
      if (is_black(f)) break;
      beta = beta * f / pf;
      prev_pdf = pf;
      d = to_world(wil, ns, ss, ts);
    }
    prev_delta = false;
    o = x;
    depth++;
  }
  return L;
}

// Render pixels [x0,x1) x [y0,y1) of a W x H image with spp samples each,
// accumulating linear radiance sums (3 doubles per pixel) into sum (indexed
// y*W + x).  Tiles of the reference's 40 x 10 layout are handed to nthreads
// workers (OfflineEngine::coreLoop, src/core/OfflineEngine.cpp:86-113).
template <class T>
void render(const World<T>& w, const CameraDesc& cd, int W, int H, int x0, int y0, int x1, int y1, int spp,
            int max_bounces, uint64_t seed, uint32_t sample_offset, uint32_t job, int nthreads, double* sum,
            PathStats* stats) {
  V3f pos(cd.position[0], cd.position[1], cd.position[2]);
  V3f at(cd.look_at[0], cd.look_at[1], cd.look_at[2]);
  V3f up(cd.up[0], cd.up[1], cd.up[2]);
  Camera cam = ref_camera(pos, at, up, cd.vfov, cd.aspect, 1.0f);
  int tw = std::max(1, (x1 - x0 + 39) / 40), th = std::max(1, (y1 - y0 + 9) / 10);
  int ntx = (x1 - x0 + tw - 1) / tw, nty = (y1 - y0 + th - 1) / th;
  int ntiles = ntx * nty;
  std::atomic<int> next{0};
  std::vector<PathStats> ps(size_t(std::max(1, nthreads)));
  auto worker = [&](int tid) {
    while (true) {
      int ti = next.fetch_add(1);
      if (ti >= ntiles) break;
      int tx = x0 + (ti % ntx) * tw, ty = y0 + (ti / ntx) * th;
      for (int y = ty; y < std::min(ty + th, y1); ++y)
        for (int x = tx; x < std::min(tx + tw, x1); ++x) {
          double acc[3] = {0, 0, 0};
          uint32_t pix = uint32_t(y) * uint32_t(W) + uint32_t(x);
          for (int si = 0; si < spp; ++si) {
            Rng rng(seed, pix, sample_offset + uint32_t(si), job);
            float ju = rng.next(), jv = rng.next();
            float u = (float(x) + ju) / float(W), v = (float(y) + jv) / float(H);
            V3f dir = ref_camera_dir(cam, u, v);
            V3<T> L = Li<T>(w, V3<T>(pos), V3<T>(dir), max_bounces, rng, stats ? &ps[size_t(tid)] : nullptr);
            acc[0] += double(L.x); acc[1] += double(L.y); acc[2] += double(L.z);
          }
          size_t o = size_t(pix) * 3;
          sum[o] += acc[0]; sum[o + 1] += acc[1]; sum[o + 2] += acc[2];
        }
    }
  };
  if (nthreads <= 1) worker(0);
  else {
    std::vector<std::thread> th_pool;
    for (int t = 0; t < nthreads; ++t) th_pool.emplace_back(worker, t);
    for (auto& t : th_pool) t.join();
  }
  if (stats)
    for (auto& p : ps) { stats->ext_rays += p.ext_rays; stats->shadow_rays += p.shadow_rays; }
}

}  // namespace neo
