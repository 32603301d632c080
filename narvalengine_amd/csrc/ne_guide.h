// ne_guide.h -- device functions of the learned directional guide shared by
// its training kernels and the guided wavefront shading stage: BF16
// conversion, hash-grid encoding, the MLP on 16x16x16 BF16 MFMA tiles, the
// vMF mixture head, density and sampling.
#pragma once
#include "ne_launch.h"
#include "ne_math.h"
#include "ne_vmf.h"

namespace ne {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kLdA = 68;  // LDS row stride (BF16) of an activation tile: conflict-free column reads
constexpr int kLdO = 49;  // LDS row stride (FP32) of an output tile
constexpr int kGuideBlock = 256;
constexpr int kWavesPerBlock = kGuideBlock / 64;

// Orders one wave's LDS writes before other lanes of the same wave read them.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- BF16 ---------------------------------------------------------------------
__host__ __device__ inline uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  uint32_t u = uint32_t(h) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// ---- hash grid ----------------------------------------------------------------
__device__ __forceinline__ uint32_t guide_cell_index(const GuideNet& g, int l, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t T = 1u << g.log2T;
  const uint32_t r1 = uint32_t(g.res[l]) + 1u;
  if ((unsigned long long)r1 * r1 * r1 <= T) return x + r1 * (y + r1 * z);  // dense level
  return (x ^ (y * 2654435761u) ^ (z * 805459861u)) & (T - 1u);
}

// Trilinear corner k (bit 0 x, bit 1 y, bit 2 z) of level l at unit position p:
// table entry index (feature f of the FP32 master at word 2 * entry + f; both
// BF16 features in table_bf[entry]) and the weight.
__device__ __forceinline__ void guide_corner(const GuideNet& g, int l, Vec3<float> p, int k, uint32_t& entry,
                                             float& w) {
  const float r = float(g.res[l]);
  const float sx = xmin(xmax(p.x, 0.0f), 1.0f) * r, sy = xmin(xmax(p.y, 0.0f), 1.0f) * r,
              sz = xmin(xmax(p.z, 0.0f), 1.0f) * r;
  // lower corner in [0, res - 1] so the upper corner stays inside the level's grid
  const float fx = xmin(floorf(sx), r - 1.0f), fy = xmin(floorf(sy), r - 1.0f), fz = xmin(floorf(sz), r - 1.0f);
  const float tx = sx - fx, ty = sy - fy, tz = sz - fz;
  const uint32_t x = uint32_t(fx) + (k & 1), y = uint32_t(fy) + ((k >> 1) & 1), z = uint32_t(fz) + ((k >> 2) & 1);
  w = ((k & 1) ? tx : 1.0f - tx) * (((k >> 1) & 1) ? ty : 1.0f - ty) * (((k >> 2) & 1) ? tz : 1.0f - tz);
  const uint32_t T = 1u << g.log2T;
  entry = uint32_t(l) * T + guide_cell_index(g, l, x, y, z);
}

// ---- MLP tile (one wave, 16 rows) -------------------------------------------------
// xs: LDS [16][kLdA] BF16 inputs of layer l; ys: LDS outputs.  Hidden layers
// write ReLU(h) as BF16 into ys; the last layer writes FP32 into out [16][kLdO].
template <int L>
__device__ __forceinline__ void guide_layer(const GuideNet& g, const uint16_t* xs, uint16_t* ys, float* out) {
  constexpr int KIN = guide_in(L), NOUT = guide_out(L);
  const int lane = __lane_id();
  const int r = lane & 15, kq = 4 * (lane >> 4);
  f32x4 acc[NOUT / 16];
#pragma unroll
  for (int nt = 0; nt < NOUT / 16; ++nt) acc[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const uint16_t* wf = g.wf + guide_w_off(L);
#pragma unroll
  for (int ks = 0; ks < KIN / 16; ++ks) {
    const bf16x4 a = *reinterpret_cast<const bf16x4*>(xs + r * kLdA + 16 * ks + kq);
#pragma unroll
    for (int nt = 0; nt < NOUT / 16; ++nt) {
      const bf16x4 b = *reinterpret_cast<const bf16x4*>(wf + (16 * nt + r) * KIN + 16 * ks + kq);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
    }
  }
  const float* bias = g.mlp + guide_b_off(L);
#pragma unroll
  for (int nt = 0; nt < NOUT / 16; ++nt) {
    const int col = 16 * nt + r;
    const float bc = bias[col];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = kq + j;
      const float h = acc[nt][j] + bc;
      if (L < kGuideLayers - 1) ys[row * kLdA + col] = f2bf(h > 0.0f ? h : 0.0f);
      else out[row * kLdO + col] = h;
    }
  }
}

// Encoding + direction of 16 queries into xs (BF16 [16][kLdA], columns 35..47 zero).
// pos/dir of row i are read through the accessor f(i, pos, dir, valid).
template <typename F>
__device__ __forceinline__ void guide_encode_tile(const GuideNet& g, uint16_t* xs, F&& f) {
  const int lane = __lane_id();
  const int row = lane & 15;
  Vec3<float> p, d;
  bool valid;
  f(row, p, d, valid);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int l = (lane >> 4) + 4 * i;
    float e0 = 0.0f, e1 = 0.0f;
    if (valid) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        uint32_t ent;
        float w;
        guide_corner(g, l, p, k, ent, w);
        const uint32_t h = g.table_bf[ent];  // 4 B per corner: features 0 and 1 as BF16
        e0 += w * bf2f(uint16_t(h & 0xFFFFu));
        e1 += w * bf2f(uint16_t(h >> 16));
      }
    }
    xs[row * kLdA + 2 * l] = f2bf(e0);
    xs[row * kLdA + 2 * l + 1] = f2bf(e1);
  }
  if (lane < 16) {
    xs[row * kLdA + 32] = f2bf(valid ? d.x : 0.0f);
    xs[row * kLdA + 33] = f2bf(valid ? d.y : 0.0f);
    xs[row * kLdA + 34] = f2bf(valid ? d.z : 0.0f);
  } else {
    for (int c = 35 + (lane >> 4) - 1; c < 64; c += 3) xs[row * kLdA + c] = 0;
  }
}

// Full forward of one tile: out [16][kLdO] raw outputs (40 used).
__device__ __forceinline__ void guide_mlp_tile(const GuideNet& g, uint16_t* xa, uint16_t* xb, float* out) {
  wave_sync();
  guide_layer<0>(g, xa, xb, out);
  wave_sync();
  guide_layer<1>(g, xb, xa, out);
  wave_sync();
  guide_layer<2>(g, xa, xb, out);
  wave_sync();
  guide_layer<3>(g, xb, xa, out);
  wave_sync();
  guide_layer<4>(g, xa, xb, out);
  wave_sync();
}

// ---- MLP on 64-row tiles (evaluation and tabulation) ------------------------------
// One wave evaluates 64 queries: four 16-row tiles share every weight fragment,
// so a layer reads its weights once per 64 queries.  K = 32 BF16 MFMA
// (v_mfma_f32_16x16x32_bf16), lane l: A row l % 16 and B column l % 16 with
// k = 8 (l / 16) .. +7 of the 32-wide k step; D column l % 16, rows
// 4 (l / 16) .. +3.  Layer 0's 48 inputs end with one K = 16 step (k = 4 (l / 16)
// .. +3, the layout of the 16-row path).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int kLd64 = 72;        // BF16 row stride of a 64-row activation tile (144 B: 16-byte aligned rows)
constexpr int kLdOut64 = 41;     // FP32 row stride of the 64-row output tile (40 used)
constexpr int kTile64Bytes = 64 * kLd64 * 2;                 // one activation tile
constexpr int kOut64Bytes = 64 * kLdOut64 * 4;               // the output tile (shares the second buffer)
constexpr int kWave64Bytes = kTile64Bytes + (kOut64Bytes > kTile64Bytes ? kOut64Bytes : kTile64Bytes);

__device__ __forceinline__ bf16x8 ld_bf16x8(const uint16_t* p) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

template <int L>
__device__ __forceinline__ void guide_layer64(const GuideNet& g, const uint16_t* xs, uint16_t* ys, float* out) {
  constexpr int KIN = guide_in(L), NOUT = guide_out(L), NT = NOUT / 16;
  const int lane = __lane_id();
  const int r = lane & 15, k8 = 8 * (lane >> 4), k4 = 4 * (lane >> 4);
  f32x4 acc[4][NT];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[rt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const uint16_t* wf = g.wf + guide_w_off(L);
#pragma unroll
  for (int ks = 0; ks < KIN / 32; ++ks) {
    bf16x8 b[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) b[nt] = ld_bf16x8(wf + (16 * nt + r) * KIN + 32 * ks + k8);
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      const bf16x8 a = ld_bf16x8(xs + (16 * rt + r) * kLd64 + 32 * ks + k8);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[rt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[nt], acc[rt][nt], 0, 0, 0);
    }
  }
  if constexpr (KIN % 32 == 16) {  // the last 16 inputs
    constexpr int k0 = KIN - 16;
    bf16x4 b[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) b[nt] = *reinterpret_cast<const bf16x4*>(wf + (16 * nt + r) * KIN + k0 + k4);
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(xs + (16 * rt + r) * kLd64 + k0 + k4);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[rt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[nt], acc[rt][nt], 0, 0, 0);
    }
  }
  const float* bias = g.mlp + guide_b_off(L);
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int col = 16 * nt + r;
    const float bc = bias[col];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = 16 * rt + k4 + j;
        const float h = acc[rt][nt][j] + bc;
        if (L < kGuideLayers - 1) ys[row * kLd64 + col] = f2bf(h > 0.0f ? h : 0.0f);
        else if (col < kGuideOutputs) out[row * kLdOut64 + col] = h;
      }
  }
}

// Encoding + direction of 64 queries (lane = row) into xs (BF16 [64][kLd64],
// columns 35..47 zero); the accessor f(row, pos, dir, valid) as for 16 rows.
template <typename F>
__device__ __forceinline__ void guide_encode_tile64(const GuideNet& g, uint16_t* xs, F&& f) {
  const int row = __lane_id();
  Vec3<float> p, d;
  bool valid;
  f(row, p, d, valid);
  uint16_t* x = xs + row * kLd64;
#pragma unroll 4
  for (int l = 0; l < kGuideLevels; ++l) {
    float e0 = 0.0f, e1 = 0.0f;
    if (valid) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        uint32_t ent;
        float w;
        guide_corner(g, l, p, k, ent, w);
        const uint32_t h = g.table_bf[ent];
        e0 += w * bf2f(uint16_t(h & 0xFFFFu));
        e1 += w * bf2f(uint16_t(h >> 16));
      }
    }
    x[2 * l] = f2bf(e0);
    x[2 * l + 1] = f2bf(e1);
  }
  x[32] = f2bf(valid ? d.x : 0.0f);
  x[33] = f2bf(valid ? d.y : 0.0f);
  x[34] = f2bf(valid ? d.z : 0.0f);
  for (int c = 35; c < 48; ++c) x[c] = 0;
}

// Full forward of a 64-row tile held in buf (kWave64Bytes): raw outputs in
// the returned [64][kLdOut64] FP32 tile.
__device__ __forceinline__ const float* guide_mlp_tile64(const GuideNet& g, unsigned char* buf) {
  uint16_t* xa = reinterpret_cast<uint16_t*>(buf);
  uint16_t* xb = reinterpret_cast<uint16_t*>(buf + kTile64Bytes);
  float* out = reinterpret_cast<float*>(buf + kTile64Bytes);
  wave_sync();
  guide_layer64<0>(g, xa, xb, out);
  wave_sync();
  guide_layer64<1>(g, xb, xa, out);
  wave_sync();
  guide_layer64<2>(g, xa, xb, out);
  wave_sync();
  guide_layer64<3>(g, xb, xa, out);
  wave_sync();
  guide_layer64<4>(g, xa, xb, out);  // the FP32 output overwrites xb, no longer read
  wave_sync();
  return out;
}

// kappa from its raw output: exp of the raw value clamped to [-10, 8]
__device__ __forceinline__ float guide_kappa(float raw) { return expf(xmin(xmax(raw, -10.0f), 8.0f)); }

// Mixture head of one output row: pi (softmax of logits), unit means, kappa.
__device__ __forceinline__ void guide_head(const float* y, float* mix) {
  float mx = y[0];
  for (int k = 1; k < kGuideLobes; ++k) mx = xmax(mx, y[5 * k]);
  float z = 0.0f;
  for (int k = 0; k < kGuideLobes; ++k) z += expf(y[5 * k] - mx);
  for (int k = 0; k < kGuideLobes; ++k) {
    const float* o = y + 5 * k;
    const float n = sqrtf(o[1] * o[1] + o[2] * o[2] + o[3] * o[3]);
    const float inv = n > 1e-12f ? 1.0f / n : 0.0f;
    mix[5 * k] = expf(o[0] - mx) / z;
    mix[5 * k + 1] = n > 1e-12f ? o[1] * inv : 0.0f;
    mix[5 * k + 2] = n > 1e-12f ? o[2] * inv : 0.0f;
    mix[5 * k + 3] = n > 1e-12f ? o[3] * inv : 1.0f;
    mix[5 * k + 4] = guide_kappa(o[4]);
  }
}

}  // namespace ne
