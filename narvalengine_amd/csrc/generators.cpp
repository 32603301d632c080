// generators.cpp -- seeded synthetic inputs for the benchmark scenes.
//
// Deterministic on any host: integer hashing (SplitMix64, Steele/Lea/Flood,
// OOPSLA 2014) seeds the permutations, and the noise itself uses only
// +, *, floor and sqrt (correctly rounded) under -ffp-contract=off.
//   * gradient noise: Perlin's improved noise (Perlin, SIGGRAPH 2002):
//     quintic fade 6t^5 - 15t^4 + 10t^3, 12 edge gradients picked by hash.
//   * cellular noise: Worley's F1 (Worley, SIGGRAPH 1996), one feature point
//     per integer cell.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ne_gpu.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
};

struct Perlin {
  int p[512];
  explicit Perlin(uint64_t seed) {
    int q[256];
    for (int i = 0; i < 256; ++i) q[i] = i;
    SplitMix64 r(seed);
    for (int i = 255; i > 0; --i) {
      int j = int(r.next() % uint64_t(i + 1));
      std::swap(q[i], q[j]);
    }
    for (int i = 0; i < 512; ++i) p[i] = q[i & 255];
  }
  static float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
  static float lerp(float t, float a, float b) { return a + t * (b - a); }
  static float grad(int h, float x, float y, float z) {
    h &= 15;
    float u = h < 8 ? x : y;
    float v = h < 4 ? y : ((h == 12 || h == 14) ? x : z);
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
  }
  float noise(float x, float y, float z) const {
    float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int X = int(fx) & 255, Y = int(fy) & 255, Z = int(fz) & 255;
    x -= fx; y -= fy; z -= fz;
    float u = fade(x), v = fade(y), w = fade(z);
    int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;
    int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
    float x1 = x - 1.0f, y1 = y - 1.0f, z1 = z - 1.0f;
    return lerp(w,
                lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x1, y, z)),
                     lerp(u, grad(p[AB], x, y1, z), grad(p[BB], x1, y1, z))),
                lerp(v, lerp(u, grad(p[AA + 1], x, y, z1), grad(p[BA + 1], x1, y, z1)),
                     lerp(u, grad(p[AB + 1], x, y1, z1), grad(p[BB + 1], x1, y1, z1))));
  }
  // 4 octaves, lacunarity 2, gain 0.5
  float fbm(float x, float y, float z) const {
    float s = 0.0f, a = 1.0f, f = 1.0f;
    for (int o = 0; o < 4; ++o) {
      s += a * noise(x * f, y * f, z * f);
      a *= 0.5f;
      f *= 2.0f;
    }
    return s;
  }
};

inline uint64_t hash3(int x, int y, int z, uint64_t seed) {
  SplitMix64 r(seed ^ (uint64_t(uint32_t(x)) * 0x8CB92BA72F3D8DD7ull) ^ (uint64_t(uint32_t(y)) * 0xD6E8FEB86659FD93ull) ^
               (uint64_t(uint32_t(z)) * 0xA0761D6478BD642Full));
  return r.next();
}

// squared distance to the nearest feature point (one per integer cell)
float worley_f1(float x, float y, float z, uint64_t seed) {
  float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
  int ix = int(fx), iy = int(fy), iz = int(fz);
  float best = 1e30f;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        uint64_t h = hash3(ix + dx, iy + dy, iz + dz, seed);
        float px = float(ix + dx) + float(h & 0xFFFFF) * (1.0f / 1048576.0f);
        float py = float(iy + dy) + float((h >> 20) & 0xFFFFF) * (1.0f / 1048576.0f);
        float pz = float(iz + dz) + float((h >> 40) & 0xFFFFF) * (1.0f / 1048576.0f);
        float ex = px - x, ey = py - y, ez = pz - z;
        float d2 = ex * ex + ey * ey + ez * ez;
        best = std::min(best, d2);
      }
  return std::sqrt(best);
}

template <class F>
void parallel_slices(int res, F&& f) {
  unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      for (int z = int(t); z < res; z += int(nt)) f(z);
    });
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

ne_status ne_gen_fbm_grid(int32_t res, uint64_t seed, float base_freq, float* out) {
  if (res <= 0 || res > 2048 || out == nullptr || !(base_freq > 0)) return NE_ERR_INVALID_ARG;
  Perlin pn(seed);
  const float inv = 1.0f / float(res);
  parallel_slices(res, [&](int z) {
    for (int y = 0; y < res; ++y)
      for (int x = 0; x < res; ++x) {
        float v = pn.fbm(float(x) * inv * base_freq, float(y) * inv * base_freq, float(z) * inv * base_freq);
        v = std::min(1.0f, std::max(0.0f, v));
        out[(size_t(z) * res + size_t(y)) * res + size_t(x)] = v;
      }
  });
  return NE_OK;
}

// density = clamp(2.5 * (E * (0.55 + 0.45 * W) - 0.3 * (0.5 + 0.5 * fbm)), 0, 1)
//   E: ellipsoid falloff 1 - |q / radii|^2 (flattened base, radii .42,.30,.42,
//      centred slightly low), W = 1 - min(1, F1(4 p)), fbm at 6 cycles/unit.
ne_status ne_gen_cumulus_grid(int32_t res, uint64_t seed, float* out) {
  if (res <= 0 || res > 2048 || out == nullptr) return NE_ERR_INVALID_ARG;
  Perlin pn(seed * 2654435761ull + 17);
  const float inv = 1.0f / float(res);
  parallel_slices(res, [&](int z) {
    for (int y = 0; y < res; ++y)
      for (int x = 0; x < res; ++x) {
        float px = float(x) * inv, py = float(y) * inv, pz = float(z) * inv;
        float qx = (px - 0.5f) / 0.42f, qy = (py - 0.45f) / 0.30f, qz = (pz - 0.5f) / 0.42f;
        if (qy < 0) qy = qy * 1.8f;  // flatter underside
        float E = 1.0f - (qx * qx + qy * qy + qz * qz);
        float v = 0.0f;
        if (E > 0.0f) {
          float W = 1.0f - std::min(1.0f, worley_f1(px * 4.0f, py * 4.0f, pz * 4.0f, seed));
          float f = 0.5f + 0.5f * pn.fbm(px * 6.0f, py * 6.0f, pz * 6.0f);
          v = 2.5f * (E * (0.55f + 0.45f * W) - 0.3f * f);
          v = std::min(1.0f, std::max(0.0f, v));
        }
        out[(size_t(z) * res + size_t(y)) * res + size_t(x)] = v;
      }
  });
  return NE_OK;
}

ne_status ne_gen_icosphere(int32_t subdiv, uint64_t seed, float amp, float* verts, uint32_t* tris,
                           uint32_t* n_vertices, uint32_t* n_triangles) {
  if (subdiv < 0 || subdiv > 9) return NE_ERR_INVALID_ARG;
  uint64_t nt = 20ull << (2 * subdiv);
  uint64_t nv = 10ull * (1ull << (2 * subdiv)) + 2;
  if (n_vertices) *n_vertices = uint32_t(nv);
  if (n_triangles) *n_triangles = uint32_t(nt);
  if (verts == nullptr || tris == nullptr) return NE_OK;
  const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
  std::vector<float> V = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                          0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
  std::vector<uint32_t> F = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6,
                             7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7,
                             9, 8, 1};
  auto norm = [&](uint32_t i) {
    float x = V[3 * i], y = V[3 * i + 1], z = V[3 * i + 2];
    float l = std::sqrt(x * x + y * y + z * z);
    V[3 * i] = x / l; V[3 * i + 1] = y / l; V[3 * i + 2] = z / l;
  };
  for (uint32_t i = 0; i < 12; ++i) norm(i);
  for (int s = 0; s < subdiv; ++s) {
    std::unordered_map<uint64_t, uint32_t> mid;
    mid.reserve(F.size());
    std::vector<uint32_t> G;
    G.reserve(F.size() * 4);
    auto midpoint = [&](uint32_t a, uint32_t b) {
      uint64_t key = a < b ? (uint64_t(a) << 32 | b) : (uint64_t(b) << 32 | a);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      uint32_t id = uint32_t(V.size() / 3);
      V.push_back(0.5f * (V[3 * a] + V[3 * b]));
      V.push_back(0.5f * (V[3 * a + 1] + V[3 * b + 1]));
      V.push_back(0.5f * (V[3 * a + 2] + V[3 * b + 2]));
      norm(id);
      mid.emplace(key, id);
      return id;
    };
    for (size_t f = 0; f < F.size(); f += 3) {
      uint32_t a = F[f], b = F[f + 1], c = F[f + 2];
      uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      uint32_t add[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      G.insert(G.end(), add, add + 12);
    }
    F.swap(G);
  }
  Perlin pn(seed);
  const size_t nvv = V.size() / 3;
  for (size_t i = 0; i < nvv; ++i) {
    float x = V[3 * i], y = V[3 * i + 1], z = V[3 * i + 2];
    float r = 1.0f + amp * pn.fbm(x * 2.0f + 7.5f, y * 2.0f + 7.5f, z * 2.0f + 7.5f);
    verts[3 * i] = x * r; verts[3 * i + 1] = y * r; verts[3 * i + 2] = z * r;
  }
  std::memcpy(tris, F.data(), F.size() * sizeof(uint32_t));
  return NE_OK;
}

}  // extern "C"
