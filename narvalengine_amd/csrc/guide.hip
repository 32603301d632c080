// guide.hip -- learned directional proposal for medium scattering vertices.
//
// Position (unit cube of the volume) -> multiresolution hash-grid encoding
// (16 levels x 2 features, trilinear) ++ incoming direction -> MLP
// 35 -> 64 -> 64 -> 64 -> 64 -> 40 (ReLU) -> mixture of 8 von Mises-Fisher
// lobes (softmax weights, normalised means, kappa = exp).  The MLP runs on
// 16x16x16 BF16 MFMA tiles (BF16 weights and layer inputs, FP32 accumulation
// and biases): a wave evaluates 16 queries per tile, passing activations
// between layers through LDS.
//
// MFMA operand layout (v_mfma_f32_16x16x16_bf16, wave64), lane l:
//   A (16 x 16, rows x k):  row l % 16, k = 4 (l / 16) .. +3
//   B (16 x 16, k x cols):  col l % 16, k = 4 (l / 16) .. +3
//   D (16 x 16):            col l % 16, rows 4 (l / 16) .. +3
//
// Training minimises the Monte-Carlo estimate of KL(target || q) with target
// proportional to Li * phase: loss = -sum_i r_i log q(w_i) / sum_i r_i over a
// batch of (position, direction, sampled direction w_i, weight r_i).  Backward
// runs layer by layer over the whole batch; weight gradients are reduced from
// per-wave partial sums in a fixed order (deterministic), hash-feature
// gradients are scattered with atomics.  Adam updates FP32 masters; BF16
// operand copies (MLP weights and hash features) are re-packed after each step.
#include "ne_guide.h"

namespace ne {

// n queries (position, direction) -> mixture (or raw outputs), 64 per wave tile.
__global__ __launch_bounds__(kGuideBlock) void k_guide_eval(GuideNet g, int n, const float* in6, float* out40,
                                                            int raw) {
  __shared__ __align__(16) unsigned char buf[kWavesPerBlock][kWave64Bytes];
  const int wv = threadIdx.x >> 6, lane = __lane_id();
  const int tiles = (n + 63) / 64;
  for (int t = blockIdx.x * kWavesPerBlock + wv; t < tiles; t += gridDim.x * kWavesPerBlock) {
    wave_sync();
    guide_encode_tile64(g, reinterpret_cast<uint16_t*>(buf[wv]), [&](int i, Vec3<float>& p, Vec3<float>& d, bool& valid) {
      const int q = t * 64 + i;
      valid = q < n;
      const float* s = in6 + 6 * size_t(valid ? q : 0);
      p = v3(s[0], s[1], s[2]);
      d = v3(s[3], s[4], s[5]);
    });
    const float* out = guide_mlp_tile64(g, buf[wv]);
    const int q = t * 64 + lane;
    if (q < n) {
      float* o = out40 + 40 * size_t(q);
      if (raw) {
        for (int c = 0; c < kGuideOutputs; ++c) o[c] = out[lane * kLdOut64 + c];
      } else {
        guide_head(&out[lane * kLdOut64], o);
      }
    }
  }
}

// Tabulated proposal: the mixture at every (cell centre, direction-bin centre)
// of the cache, written to cache.mix (40 floats per entry).
__global__ __launch_bounds__(kGuideBlock) void k_guide_bake(GuideNet g, GuideCache c, float* mix, int n) {
  __shared__ __align__(16) unsigned char buf[kWavesPerBlock][kWave64Bytes];
  const int wv = threadIdx.x >> 6, lane = __lane_id();
  const int nb = c.nd * c.nd;
  const int tiles = (n + 63) / 64;
  for (int t = blockIdx.x * kWavesPerBlock + wv; t < tiles; t += gridDim.x * kWavesPerBlock) {
    wave_sync();
    guide_encode_tile64(g, reinterpret_cast<uint16_t*>(buf[wv]), [&](int i, Vec3<float>& p, Vec3<float>& d, bool& valid) {
      const int q = t * 64 + i;
      valid = q < n;
      const int e = valid ? q : 0;
      const int cell = e / nb, bin = e % nb;
      const int cx = cell % c.nc, cy = (cell / c.nc) % c.nc, cz = cell / (c.nc * c.nc);
      const float inv = 1.0f / float(c.nc);
      p = v3((float(cx) + 0.5f) * inv, (float(cy) + 0.5f) * inv, (float(cz) + 0.5f) * inv);
      const float invd = 1.0f / float(c.nd);
      oct_decode((float(bin % c.nd) + 0.5f) * invd, (float(bin / c.nd) + 0.5f) * invd, d.x, d.y, d.z);
    });
    const float* out = guide_mlp_tile64(g, buf[wv]);
    const int q = t * 64 + lane;
    if (q < n) guide_head(&out[lane * kLdOut64], mix + 40 * size_t(q));
  }
}

// ---- training: forward + loss gradient --------------------------------------------
// Per tile: layer inputs stored for backward (act[l][row][64] BF16), loss
// gradient w.r.t. the 40 raw outputs into dy (BF16, columns 40..47 zero).
__global__ __launch_bounds__(kGuideBlock) void k_guide_train_fwd(GuideNet g, GuideBatch b, GuideTrain t) {
  __shared__ uint16_t xa[kWavesPerBlock][16 * kLdA];
  __shared__ uint16_t xb[kWavesPerBlock][16 * kLdA];
  __shared__ float out[kWavesPerBlock][16 * kLdO];
  const int wv = threadIdx.x >> 6, lane = __lane_id();
  const uint32_t n = min(*b.count, b.cap);
  const uint32_t tiles = (n + 15) / 16;
  const float* S = b.s;
  const uint32_t cap = b.cap;
  float loss = 0.0f, wsum = 0.0f;
  for (uint32_t tt = blockIdx.x * kWavesPerBlock + wv; tt < tiles; tt += gridDim.x * kWavesPerBlock) {
    wave_sync();
    guide_encode_tile(g, xa[wv], [&](int i, Vec3<float>& p, Vec3<float>& d, bool& valid) {
      const uint32_t q = tt * 16 + uint32_t(i);
      valid = q < n;
      const uint32_t k = valid ? q : 0u;
      p = v3(S[k], S[cap + k], S[2 * cap + k]);
      d = v3(S[3 * cap + k], S[4 * cap + k], S[5 * cap + k]);
    });
    // keep every layer input for backward: act[l][row][c]
    uint16_t* bufs[2] = {xa[wv], xb[wv]};
#define NE_STORE_ACT(L, SRC)                                                                 \
    for (int e = lane; e < 16 * 64; e += 64) {                                               \
      const int row = e >> 6, c = e & 63;                                                    \
      const uint32_t q = tt * 16 + uint32_t(row);                                            \
      if (q < n) t.act[(size_t(L) * cap + q) * kGuideWidth + c] = c < guide_in(L) ? (SRC)[row * kLdA + c] : 0; \
    }
    wave_sync();
    NE_STORE_ACT(0, bufs[0]);
    guide_layer<0>(g, xa[wv], xb[wv], out[wv]);
    wave_sync();
    NE_STORE_ACT(1, bufs[1]);
    guide_layer<1>(g, xb[wv], xa[wv], out[wv]);
    wave_sync();
    NE_STORE_ACT(2, bufs[0]);
    guide_layer<2>(g, xa[wv], xb[wv], out[wv]);
    wave_sync();
    NE_STORE_ACT(3, bufs[1]);
    guide_layer<3>(g, xb[wv], xa[wv], out[wv]);
    wave_sync();
    NE_STORE_ACT(4, bufs[0]);
    guide_layer<4>(g, xa[wv], xb[wv], out[wv]);
    wave_sync();
#undef NE_STORE_ACT
    if (lane < 16) {
      const uint32_t q = tt * 16 + uint32_t(lane);
      const float* y = &out[wv][lane * kLdO];
      float dy[kGuideOutputs];
      for (int c = 0; c < kGuideOutputs; ++c) dy[c] = 0.0f;
      if (q < n) {
        const Vec3<float> w = v3(S[6 * cap + q], S[7 * cap + q], S[8 * cap + q]);
        const float r = S[9 * cap + q];
        float mx = y[0];
        for (int k = 1; k < kGuideLobes; ++k) mx = xmax(mx, y[5 * k]);
        float z = 0.0f;
        for (int k = 0; k < kGuideLobes; ++k) z += expf(y[5 * k] - mx);
        float pi[kGuideLobes], v[kGuideLobes], qsum = 0.0f;
        for (int k = 0; k < kGuideLobes; ++k) {
          const float* o = y + 5 * k;
          pi[k] = expf(o[0] - mx) / z;
          const float nm = xmax(sqrtf(o[1] * o[1] + o[2] * o[2] + o[3] * o[3]), 1e-12f);
          const float c = (o[1] * w.x + o[2] * w.y + o[3] * w.z) / nm;
          const float kap = guide_kappa(o[4]);
          v[k] = vmf_pdf(kap, c);
          qsum += pi[k] * v[k];
        }
        qsum = xmax(qsum, 1e-30f);
        loss += -r * logf(qsum);
        wsum += r;
        const float s = -r / qsum;  // d(-r log q) / dq
        for (int k = 0; k < kGuideLobes; ++k) {
          const float* o = y + 5 * k;
          dy[5 * k] = s * pi[k] * (v[k] - qsum);
          const float nm = xmax(sqrtf(o[1] * o[1] + o[2] * o[2] + o[3] * o[3]), 1e-12f);
          const float mu[3] = {o[1] / nm, o[2] / nm, o[3] / nm};
          const float c = mu[0] * w.x + mu[1] * w.y + mu[2] * w.z;
          const float kap = guide_kappa(o[4]);
          const float pv = s * pi[k] * v[k];
          const float wk[3] = {w.x, w.y, w.z};
          for (int a = 0; a < 3; ++a) dy[5 * k + 1 + a] = pv * kap * (wk[a] - c * mu[a]) / nm;
          dy[5 * k + 4] = (o[4] < 8.0f && o[4] > -10.0f) ? pv * kap * vmf_dlog_dkappa(kap, c) : 0.0f;
        }
        for (int c = 0; c < 48; ++c) t.dy[size_t(q) * kGuideWidth + c] = f2bf(c < kGuideOutputs ? dy[c] : 0.0f);
        for (int c = 48; c < 64; ++c) t.dy[size_t(q) * kGuideWidth + c] = 0;
      }
    }
  }
  // loss statistics: one atomic pair per wave
  for (int o = 32; o > 0; o >>= 1) {
    loss += __shfl_down(loss, o);
    wsum += __shfl_down(wsum, o);
  }
  if (lane == 0) {
    atomicAdd(&t.loss[0], loss);
    atomicAdd(&t.loss[1], wsum);
  }
}

// ---- training: backward of layer L ------------------------------------------------
// dy: gradient w.r.t. this layer's (pre-ReLU for hidden, raw for the last) output,
// BF16 [n][64].  Writes dx = dy W^T masked by the ReLU of this layer's input
// (which is the previous layer's output) into dx (BF16) or, for L = 0, the
// encoding gradient into denc (FP32); accumulates this wave's dW / db partial.
template <int L>
__global__ __launch_bounds__(kGuideBlock) void k_guide_train_bwd(GuideNet g, GuideBatch b, GuideTrain t,
                                                                 const uint16_t* dy, uint16_t* dx) {
  constexpr int KIN = guide_in(L), NOUT = guide_out(L);
  const float scale = t.loss[1] > 0.0f ? 1.0f / t.loss[1] : 0.0f;  // loss normalised by sum r
  const int lane = __lane_id();
  const int r = lane & 15, kq = 4 * (lane >> 4);
  const int wid = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  const uint32_t n = min(*b.count, b.cap);
  const uint32_t tiles = (n + 15) / 16;
  const uint32_t cap = b.cap;
  const uint16_t* X = t.act + size_t(L) * cap * kGuideWidth;
  const uint16_t* wb = g.wb + guide_w_off(L);
  f32x4 dw[KIN / 16][NOUT / 16];
#pragma unroll
  for (int mt = 0; mt < KIN / 16; ++mt)
#pragma unroll
    for (int nt = 0; nt < NOUT / 16; ++nt) dw[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float db[NOUT / 16];
#pragma unroll
  for (int nt = 0; nt < NOUT / 16; ++nt) db[nt] = 0.0f;
  for (uint32_t tt = uint32_t(wid); tt < tiles; tt += uint32_t(t.waves)) {
    const uint32_t row0 = tt * 16;
    // dx = dy W^T: A = dy rows (k = output index), B = W^T (k = output, n = input)
    f32x4 ax[KIN / 16];
#pragma unroll
    for (int mt = 0; mt < KIN / 16; ++mt) ax[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t ra = row0 + uint32_t(r);
#pragma unroll
    for (int ks = 0; ks < NOUT / 16; ++ks) {
      bf16x4 a = bf16x4{0, 0, 0, 0};
      if (ra < n) a = *reinterpret_cast<const bf16x4*>(dy + size_t(ra) * kGuideWidth + 16 * ks + kq);
#pragma unroll
      for (int nt = 0; nt < KIN / 16; ++nt) {
        const bf16x4 bb = *reinterpret_cast<const bf16x4*>(wb + (16 * nt + r) * NOUT + 16 * ks + kq);
        ax[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, ax[nt], 0, 0, 0);
      }
    }
#pragma unroll
    for (int nt = 0; nt < KIN / 16; ++nt) {
      const int col = 16 * nt + r;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t row = row0 + uint32_t(kq + j);
        if (row >= n) continue;
        if (L == 0) {
          if (col < kGuideEnc) t.denc[size_t(row) * kGuideEnc + col] = ax[nt][j];
        } else {
          const float xin = bf2f(X[size_t(row) * kGuideWidth + col]);  // ReLU output of layer L-1
          dx[size_t(row) * kGuideWidth + col] = f2bf(xin > 0.0f ? ax[nt][j] : 0.0f);
        }
      }
    }
    // dW += X^T dy: A = X^T (m = input, k = row), B = dy (k = row, n = output)
#pragma unroll
    for (int mt = 0; mt < KIN / 16; ++mt) {
      bf16x4 a;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t row = row0 + uint32_t(kq + j);
        a[j] = row < n ? short(X[size_t(row) * kGuideWidth + 16 * mt + r]) : short(0);
      }
#pragma unroll
      for (int nt = 0; nt < NOUT / 16; ++nt) {
        bf16x4 bb;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t row = row0 + uint32_t(kq + j);
          bb[j] = row < n ? short(dy[size_t(row) * kGuideWidth + 16 * nt + r]) : short(0);
        }
        dw[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, dw[mt][nt], 0, 0, 0);
      }
    }
#pragma unroll
    for (int nt = 0; nt < NOUT / 16; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t row = row0 + uint32_t(kq + j);
        if (row < n) db[nt] += bf2f(dy[size_t(row) * kGuideWidth + 16 * nt + r]);
      }
  }
  // this wave's partial: dW[in][out] (lane: rows kq..kq+3 of each m-tile, column r of each n-tile)
  float* part = t.partial + size_t(wid) * kGuideParams;
#pragma unroll
  for (int mt = 0; mt < KIN / 16; ++mt)
#pragma unroll
    for (int nt = 0; nt < NOUT / 16; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        part[guide_w_off(L) + (16 * mt + kq + j) * NOUT + 16 * nt + r] = dw[mt][nt][j] * scale;
#pragma unroll
  for (int nt = 0; nt < NOUT / 16; ++nt) {
    float s = db[nt];
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    if (lane < 16) part[guide_b_off(L) + 16 * nt + r] = s * scale;
  }
}

// mlp_grad[p] = sum over waves of the partials, in a fixed order
// (deterministic): a block takes 64 parameters, its 4 thread groups sum the
// waves w = j mod 4 each, and the 4 group sums are added in order.
__global__ __launch_bounds__(256) void k_guide_reduce(GuideNet g, GuideTrain t) {
  __shared__ float part[4][64];
  const int j = int(threadIdx.x >> 6), col = int(threadIdx.x & 63);
  const int p = int(blockIdx.x) * 64 + col;
  float s = 0.0f;
  if (p < kGuideParams)
    for (int w = j; w < t.waves; w += 4) s += t.partial[size_t(w) * kGuideParams + p];
  part[j][col] = s;
  __syncthreads();
  if (j == 0 && p < kGuideParams) g.mlp_grad[p] = ((part[0][col] + part[1][col]) + part[2][col]) + part[3][col];
}

// hash-feature gradients: scatter denc * corner weight (scaled) into
// table_grad.  One thread per sample, levels in turn.  On the coarse levels
// (resolution <= 64) the samples of a wave mostly share corners, so the wave
// first merges equal entries: the lowest pending lane's entry is taken, the
// lanes holding it are summed across the wave and one lane adds the sum (up
// to 4 entries per corner); whatever is left, and every finer level, adds
// directly.
__global__ __launch_bounds__(256) void k_guide_table_grad(GuideNet g, GuideBatch b, GuideTrain t) {
  const float scale = t.loss[1] > 0.0f ? 1.0f / t.loss[1] : 0.0f;
  const uint32_t n = min(*b.count, b.cap);
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = q < n;  // no early exit: every lane takes part in the wave merges
  const int lane = __lane_id();
  const uint32_t qq = valid ? q : 0u;
  const Vec3<float> p = v3(b.s[qq], b.s[b.cap + qq], b.s[2 * b.cap + qq]);
  for (int l = 0; l < kGuideLevels; ++l) {
    const float d0 = valid ? t.denc[size_t(qq) * kGuideEnc + 2 * l] * scale : 0.0f;
    const float d1 = valid ? t.denc[size_t(qq) * kGuideEnc + 2 * l + 1] * scale : 0.0f;
    const bool coarse = g.res[l] <= 64;  // the same for every lane
    for (int k = 0; k < 8; ++k) {
      uint32_t ent;
      float w;
      guide_corner(g, l, p, k, ent, w);
      const float v0 = w * d0, v1 = w * d1;
      bool pend = valid && (v0 != 0.0f || v1 != 0.0f);
      if (coarse) {
        for (int it = 0; it < 4; ++it) {
          const uint64_t m = __ballot(pend);
          if (m == 0) break;
          const int leader = __ffsll((long long)m) - 1;
          const uint32_t key = uint32_t(__shfl(int(ent), leader));
          const bool mine = pend && ent == key;
          float s0 = mine ? v0 : 0.0f, s1 = mine ? v1 : 0.0f;
          for (int o = 32; o > 0; o >>= 1) {
            s0 += __shfl_xor(s0, o);
            s1 += __shfl_xor(s1, o);
          }
          if (lane == leader) {
            atomicAdd(&g.table_grad[2 * size_t(key)], s0);
            atomicAdd(&g.table_grad[2 * size_t(key) + 1], s1);
          }
          pend = pend && !mine;
        }
      }
      if (pend) {
        atomicAdd(&g.table_grad[2 * size_t(ent)], v0);
        atomicAdd(&g.table_grad[2 * size_t(ent) + 1], v1);
      }
    }
  }
}

// BF16 copy of the hash features that the encoder reads (after every change
// of the FP32 masters)
__global__ void k_guide_pack_table(GuideNet g, size_t entries) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= entries) return;
  const float2 v = reinterpret_cast<const float2*>(g.table)[i];
  g.table_bf[i] = uint32_t(f2bf(v.x)) | (uint32_t(f2bf(v.y)) << 16);
}

__global__ void k_guide_adam(float* prm, float* grad, float* m, float* v, size_t n, float lr, float b1, float b2,
                             float eps, float c1, float c2) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gr = grad[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gr;
  const float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
  m[i] = mi;
  v[i] = vi;
  prm[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  grad[i] = 0.0f;
}

__global__ void k_guide_pack(GuideNet g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kGuideWeights) return;
  int l = 0;
  while (l + 1 < kGuideLayers && i >= guide_w_off(l + 1)) ++l;
  const int local = i - guide_w_off(l), in = local / guide_out(l), out = local % guide_out(l);
  const uint16_t h = f2bf(g.mlp[i]);
  g.wb[i] = h;                                                   // [in][out]
  g.wf[guide_w_off(l) + out * guide_in(l) + in] = h;             // [out][in]
}

// Initial parameters: hash features U(-1e-4, 1e-4); weights U(-s, s) with
// s = sqrt(6 / fan_in) over the used inputs (He-uniform), padding zero; biases 0.
__device__ __forceinline__ float guide_u01(uint64_t seed, uint32_t stream, uint32_t i) {
  uint32_t c[4] = {i, stream, 0x6775u, 0u};
  philox10(c, uint32_t(seed), uint32_t(seed >> 32));
  return float(c[0] >> 8) * (1.0f / 16777216.0f);
}
__global__ void k_guide_init(GuideNet g, uint64_t seed, size_t table_n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < table_n) {
    g.table[i] = (2.0f * guide_u01(seed, 1u, uint32_t(i)) - 1.0f) * 1e-4f;
    g.table_grad[i] = 0.0f; g.table_m[i] = 0.0f; g.table_v[i] = 0.0f;
  }
  if (i < size_t(kGuideParams)) {
    float val = 0.0f;
    if (i < size_t(kGuideWeights)) {
      int l = 0;
      while (l + 1 < kGuideLayers && int(i) >= guide_w_off(l + 1)) ++l;
      const int local = int(i) - guide_w_off(l), in = local / guide_out(l), out = local % guide_out(l);
      const int used_in = l == 0 ? kGuideInputs : 64, used_out = l == kGuideLayers - 1 ? kGuideOutputs : 64;
      if (in < used_in && out < used_out) {
        const float s = sqrtf(6.0f / float(used_in));
        val = (2.0f * guide_u01(seed, 2u, uint32_t(i)) - 1.0f) * s;
      }
    }
    g.mlp[i] = val;
    g.mlp_grad[i] = 0.0f; g.mlp_m[i] = 0.0f; g.mlp_v[i] = 0.0f;
  }
}

// ---- launchers ------------------------------------------------------------------------
static size_t table_words(const GuideNet& g) { return size_t(kGuideLevels) * (size_t(1) << g.log2T) * kGuideFeat; }

hipError_t launch_guide_init(const GuideNet& g, uint64_t seed, hipStream_t st) {
  const size_t n = table_words(g);
  hipLaunchKernelGGL(k_guide_init, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, st, g, seed, n);
  return launch_guide_pack(g, st);
}

hipError_t launch_guide_pack(const GuideNet& g, hipStream_t st) {
  hipLaunchKernelGGL(k_guide_pack, dim3((kGuideWeights + 255) / 256), dim3(256), 0, st, g);
  const size_t entries = table_words(g) / kGuideFeat;
  hipLaunchKernelGGL(k_guide_pack_table, dim3(uint32_t((entries + 255) / 256)), dim3(256), 0, st, g, entries);
  return hipGetLastError();
}

hipError_t launch_guide_eval(const GuideNet& g, int n, const float* in6, float* out40, hipStream_t st) {
  const int tiles = (n + 63) / 64;
  const int blocks = std::max(1, std::min(4096, (tiles + kWavesPerBlock - 1) / kWavesPerBlock));
  hipLaunchKernelGGL(k_guide_eval, dim3(blocks), dim3(kGuideBlock), 0, st, g, n, in6, out40, 0);
  return hipGetLastError();
}

hipError_t launch_guide_raw(const GuideNet& g, int n, const float* in6, float* out40, hipStream_t st) {
  const int tiles = (n + 63) / 64;
  const int blocks = std::max(1, std::min(4096, (tiles + kWavesPerBlock - 1) / kWavesPerBlock));
  hipLaunchKernelGGL(k_guide_eval, dim3(blocks), dim3(kGuideBlock), 0, st, g, n, in6, out40, 1);
  return hipGetLastError();
}

hipError_t launch_guide_bake(const GuideNet& g, const GuideCache& c, float* mix, hipStream_t st) {
  const int n = c.nc * c.nc * c.nc * c.nd * c.nd;
  const int tiles = (n + 63) / 64;
  const int blocks = std::max(1, std::min(8192, (tiles + kWavesPerBlock - 1) / kWavesPerBlock));
  hipLaunchKernelGGL(k_guide_bake, dim3(blocks), dim3(kGuideBlock), 0, st, g, c, mix, n);
  return hipGetLastError();
}

__global__ void k_guide_zero_loss(GuideTrain t) {
  if (threadIdx.x < 2) t.loss[threadIdx.x] = 0.0f;
}

// Gradients of the weight-normalised loss -sum r log q / sum r (the scale
// 1 / sum r is read on the device after the forward pass).
hipError_t launch_guide_grad(const GuideNet& g, const GuideBatch& b, const GuideTrain& t, hipStream_t st) {
  const dim3 blk(kGuideBlock);
  const uint32_t blocks = uint32_t(t.waves / kWavesPerBlock);
  hipLaunchKernelGGL(k_guide_zero_loss, dim3(1), dim3(64), 0, st, t);
  hipLaunchKernelGGL(k_guide_train_fwd, dim3(blocks), blk, 0, st, g, b, t);
  hipLaunchKernelGGL((k_guide_train_bwd<4>), dim3(blocks), blk, 0, st, g, b, t, t.dy, t.dy2);
  hipLaunchKernelGGL((k_guide_train_bwd<3>), dim3(blocks), blk, 0, st, g, b, t, t.dy2, t.dy);
  hipLaunchKernelGGL((k_guide_train_bwd<2>), dim3(blocks), blk, 0, st, g, b, t, t.dy, t.dy2);
  hipLaunchKernelGGL((k_guide_train_bwd<1>), dim3(blocks), blk, 0, st, g, b, t, t.dy2, t.dy);
  hipLaunchKernelGGL((k_guide_train_bwd<0>), dim3(blocks), blk, 0, st, g, b, t, t.dy, t.dy2);
  hipLaunchKernelGGL(k_guide_reduce, dim3((kGuideParams + 63) / 64), dim3(256), 0, st, g, t);
  hipLaunchKernelGGL(k_guide_table_grad, dim3((b.cap + 255) / 256), dim3(256), 0, st, g, b, t);
  return hipGetLastError();
}

hipError_t launch_guide_adam(const GuideNet& g, float lr_table, float lr_mlp, float beta1, float beta2, float eps,
                             int step, hipStream_t st) {
  const float c1 = 1.0f - powf(beta1, float(step)), c2 = 1.0f - powf(beta2, float(step));
  const size_t nt = table_words(g);
  hipLaunchKernelGGL(k_guide_adam, dim3(uint32_t((nt + 255) / 256)), dim3(256), 0, st, g.table, g.table_grad,
                     g.table_m, g.table_v, nt, lr_table, beta1, beta2, eps, c1, c2);
  hipLaunchKernelGGL(k_guide_adam, dim3((kGuideParams + 255) / 256), dim3(256), 0, st, g.mlp, g.mlp_grad, g.mlp_m,
                     g.mlp_v, size_t(kGuideParams), lr_mlp, beta1, beta2, eps, c1, c2);
  return launch_guide_pack(g, st);
}

}  // namespace ne
