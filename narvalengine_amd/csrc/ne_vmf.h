// ne_vmf.h -- von Mises-Fisher mixtures (density, sampling) and the
// tabulated guide proposal that the path tracers sample at medium vertices.
#pragma once
#include "ne_math.h"

namespace ne {

constexpr int kVmfLobes = 8;  // lobes of a mixture: 5 floats each (weight, unit mean xyz, kappa)

// ---- vMF mixture ------------------------------------------------------------------
// vMF density kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1)), stable for
// large kappa; the normaliser equals kappa / (4 pi sinh kappa) exp(kappa).
__device__ __forceinline__ float vmf_pdf(float kappa, float cosv) {
  return kappa / (float(kTWO_PI) * -expm1f(-2.0f * kappa)) * expf(kappa * (cosv - 1.0f));
}
// d log vMF / d kappa = c - L(kappa), L the Langevin function coth k - 1/k
__device__ __forceinline__ float vmf_dlog_dkappa(float kappa, float cosv) {
  const float L = kappa < 1e-2f ? kappa / 3.0f - kappa * kappa * kappa / 45.0f : 1.0f / tanhf(kappa) - 1.0f / kappa;
  return cosv - L;
}

// vMF sample around mu (inverse CDF of cos theta in the stable form).
// cos theta = 1 + log(u + (1 - u) exp(-2 kappa)) / kappa, written with
// log1p / expm1 so it stays exact as kappa -> 0 (uniform sphere).
__device__ __forceinline__ Vec3<float> vmf_sample(Vec3<float> mu, float kappa, float u0, float u1) {
  const float c = 1.0f + log1pf((1.0f - u0) * expm1f(-2.0f * kappa)) / kappa;
  const float cc = xmin(1.0f, xmax(-1.0f, c));
  const float sn = sqrtf(xmax(0.0f, 1.0f - cc * cc));
  const float ph = float(kTWO_PI) * u1;
  Vec3<float> ss, ts;
  onb(mu, ss, ts);
  return to_world(v3(sn * cosf(ph), sn * sinf(ph), cc), mu, ss, ts);
}

// Mixture density at w; mix: 8 lobes of (weight, unit mean xyz, kappa).
__device__ __forceinline__ float vmf_mixture_pdf(const float* mix, Vec3<float> w) {
  float q = 0.0f;
  for (int k = 0; k < kVmfLobes; ++k) {
    const float* m = mix + 5 * k;
    q += m[0] * vmf_pdf(m[4], m[1] * w.x + m[2] * w.y + m[3] * w.z);
  }
  return q;
}

// Lobe chosen by u (inverse CDF over the weights), then a vMF sample.
__device__ __forceinline__ Vec3<float> vmf_mixture_sample(const float* mix, float u, float u0, float u1) {
  int k = 0;
  float acc = mix[0];
  while (k < kVmfLobes - 1 && u >= acc) {
    ++k;
    acc += mix[5 * k];
  }
  const float* m = mix + 5 * k;
  return vmf_sample(v3(m[1], m[2], m[3]), m[4], u0, u1);
}

// ---- tabulated proposal ----------------------------------------------------------
// The guide's mixture at the centre of each of nc^3 cells of the volume's unit
// cube and of each of nd x nd octahedral bins of the incoming direction: the
// network is evaluated once per entry after every training step, and a
// scattering vertex reads the entry of its cell and direction.
struct GuideCache {
  const float* mix;  // (nc^3 * nd^2) entries of 40 floats, cell-major, bins inner
  int32_t nc, nd;
};
// How a render uses it: vertices at depth < max_depth sample the one-sample
// MIS mixture (1 - prob) HG + prob proposal; cache.mix == nullptr: plain HG.
struct GuideUse {
  GuideCache cache;
  int32_t max_depth;
  float prob;
};

// Octahedral map of a unit direction onto [0, 1]^2 (upper hemisphere in the
// inner diamond, lower one folded onto the corners).
__host__ __device__ __forceinline__ void oct_encode(float x, float y, float z, float& u, float& v) {
  const float inv = 1.0f / (fabsf(x) + fabsf(y) + fabsf(z));
  float px = x * inv, py = y * inv;
  if (z < 0.0f) {
    const float ox = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
    const float oy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
    px = ox;
    py = oy;
  }
  u = px * 0.5f + 0.5f;
  v = py * 0.5f + 0.5f;
}
__host__ __device__ __forceinline__ void oct_decode(float u, float v, float& x, float& y, float& z) {
  float px = 2.0f * u - 1.0f, py = 2.0f * v - 1.0f;
  const float pz = 1.0f - fabsf(px) - fabsf(py);
  if (pz < 0.0f) {
    const float ox = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
    const float oy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
    px = ox;
    py = oy;
  }
  const float n = sqrtf(px * px + py * py + pz * pz);
  x = px / n;
  y = py / n;
  z = pz / n;
}
__host__ __device__ __forceinline__ int oct_bin(float x, float y, float z, int nd) {
  float u, v;
  oct_encode(x, y, z, u, v);
  const int bx = min(max(int(u * float(nd)), 0), nd - 1), by = min(max(int(v * float(nd)), 0), nd - 1);
  return by * nd + bx;
}
// Entry of the unit-cube position pu (clamped into the cube) and direction d.
__device__ __forceinline__ const float* guide_cache_at(const GuideCache& c, Vec3<float> pu, Vec3<float> d) {
  const float fn = float(c.nc);
  const int cx = min(max(int(pu.x * fn), 0), c.nc - 1), cy = min(max(int(pu.y * fn), 0), c.nc - 1),
            cz = min(max(int(pu.z * fn), 0), c.nc - 1);
  const size_t cell = (size_t(cz) * size_t(c.nc) + size_t(cy)) * size_t(c.nc) + size_t(cx);
  return c.mix + 40 * (cell * size_t(c.nd * c.nd) + size_t(oct_bin(d.x, d.y, d.z, c.nd)));
}

}  // namespace ne
