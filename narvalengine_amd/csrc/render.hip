// render.hip -- path-tracing kernels for gfx950 and their launchers.
//
// k_render: one thread per (pixel, chunk of S samples).  A wavefront owns an
// 8x8 pixel tile (adjacent lanes trace adjacent pixels, so BVH nodes and grid
// cells are shared through L1/L2), blocks are remapped so each XCD walks a
// contiguous band of tiles (its own 4 MB L2 sees neighbouring tiles), and
// every lane keeps its BVH short stack in LDS.  Per-pixel sums are FP64 in
// registers and written without atomics, so the result is bit-reproducible
// for any launch shape: chunk partials are summed in chunk order by
// k_reduce_partials.
#include <type_traits>

#include "ne_launch.h"
#include "ne_trace.h"

namespace ne {

constexpr int kNumXcd = 8;

// Logical block index with XCD-contiguous ranges.  The grid is padded to a
// multiple of 8 blocks; hardware places block b on XCD b % 8.
__device__ __forceinline__ uint32_t xcd_logical_block(uint32_t nreal) {
  uint32_t nb = gridDim.x, b = blockIdx.x;
  uint32_t per = nb / kNumXcd;
  uint32_t lb = (b % kNumXcd) * per + b / kNumXcd;
  (void)nreal;
  return lb;
}

// SURF = false: a volume lit only by the sun and the environment (no surface
// code); GUIDED: medium vertices sample the guide's tabulated proposal.
template <typename R, bool C, int BLOCK, int WAVES = 0, bool SURF = true, bool GUIDED = false>
__global__ __launch_bounds__(BLOCK, WAVES) void k_render(DevScene s, DevCamera cam, RenderArgs a,
                                                          uint32_t nreal_blocks) {
  __shared__ uint32_t lds_stack[kShortStack * BLOCK];
  uint32_t lb = xcd_logical_block(nreal_blocks);
  if (lb >= nreal_blocks) return;
  const int lane = threadIdx.x & 63;
  const uint32_t gw = lb * (BLOCK / 64) + (threadIdx.x >> 6);
  const uint32_t ntiles = uint32_t(a.ntx) * uint32_t(a.nty);
  const uint32_t tile = gw % ntiles, chunk = gw / ntiles;
  if (chunk >= uint32_t(a.nchunks)) return;
  const int px = a.x0 + int(tile % a.ntx) * 8 + (lane & 7);
  const int py = a.y0 + int(tile / a.ntx) * 8 + (lane >> 3);
  if (px >= a.x1 || py >= a.y1) return;
  uint32_t* lds = lds_stack + threadIdx.x;
  Count<C> cnt;
  const uint32_t pix = uint32_t(py) * uint32_t(a.W) + uint32_t(px);
  const int s0 = int(chunk) * a.S;
  const int s1 = s0 + a.S < a.spp ? s0 + a.S : a.spp;
  Vec3<float> pos = v3f<float>(cam.pos), ll = v3f<float>(cam.lower_left);
  Vec3<float> hz = v3f<float>(cam.horizontal), vt = v3f<float>(cam.vertical);
  double acc0 = 0, acc1 = 0, acc2 = 0;
  for (int si = s0; si < s1; ++si) {
    PathRng rng;
    rng.init(a.seed, pix, a.sample_offset + uint32_t(si), a.job);
    float ju = rng.next(), jv = rng.next();
    float u = (float(px) + ju) / float(a.W), v = (float(py) + jv) / float(a.H);
    Vec3<float> dir = -normalize(ll + u * hz + v * vt - pos);
    Vec3<R> L = trace_path<R, C, SURF, GUIDED>(s, vcast<R>(pos), vcast<R>(dir), a.max_bounces, rng, lds, BLOCK, cnt,
                                               &a.guide);
    acc0 += double(L.x); acc1 += double(L.y); acc2 += double(L.z);
  }
  if (a.nchunks == 1) {
    double* o = a.out + size_t(pix) * 3;
    o[0] += acc0; o[1] += acc1; o[2] += acc2;
  } else {
    const int rw = a.x1 - a.x0, rh = a.y1 - a.y0;
    size_t q = (size_t(chunk) * size_t(rw) * size_t(rh) + size_t(py - a.y0) * size_t(rw) + size_t(px - a.x0)) * 3;
    a.out[q] = acc0; a.out[q + 1] = acc1; a.out[q + 2] = acc2;
  }
  if constexpr (C) {
    atomicAdd(&a.counters->paths, (unsigned long long)(s1 - s0));
    atomicAdd(&a.counters->ext_rays, (unsigned long long)cnt.n_ext);
    atomicAdd(&a.counters->shadow_rays, (unsigned long long)cnt.n_shadow);
    atomicAdd(&a.counters->density_lookups, (unsigned long long)cnt.n_lookup);
    atomicAdd(&a.counters->majorant_cells, (unsigned long long)cnt.n_cell);
    atomicAdd(&a.counters->bvh_nodes, (unsigned long long)cnt.n_node);
    atomicAdd(&a.counters->prim_tests, (unsigned long long)cnt.n_prim);
  }
}

__global__ void k_reduce_partials(RenderArgs a, double* film) {
  const int rw = a.x1 - a.x0, rh = a.y1 - a.y0;
  const size_t np = size_t(rw) * size_t(rh);
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int c = 0; c < a.nchunks; ++c) {
    const double* p = a.out + (size_t(c) * np + i) * 3;
    s0 += p[0]; s1 += p[1]; s2 += p[2];
  }
  int px = a.x0 + int(i % size_t(rw)), py = a.y0 + int(i / size_t(rw));
  double* o = film + (size_t(py) * size_t(a.W) + size_t(px)) * 3;
  o[0] += s0; o[1] += s1; o[2] += s2;
}

// OfflineEngine::postProcessing (src/core/OfflineEngine.cpp:27-52):
// clamp(pow(1 - exp(-0.5 c), 1/2.2), 0, 1) per channel.
__device__ __forceinline__ float post_process(double sum, float inv_spp) {
  float c = float(sum * double(inv_spp));
  float m = 1.0f - expf(-c * 0.5f);
  m = powf(m, 1.0f / 2.2f);
  return xmin(xmax(m, 0.0f), 1.0f);
}
__global__ void k_tonemap(const double* film, int n, float inv_spp, float* pixels) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * 3) return;
  pixels[i] = post_process(film[i], inv_spp);
}
// The same for one pixel rectangle, packed: out[(y * rw + x) * 3 + ch].
__global__ void k_tonemap_rect(const double* film, int W, int x0, int y0, int rw, int rh, float inv_spp, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rw * rh * 3) return;
  const int ch = i % 3, p = i / 3, x = p % rw, y = p / rw;
  out[i] = post_process(film[(size_t(y0 + y) * W + size_t(x0 + x)) * 3 + ch], inv_spp);
}

// Zeroes the film sums of the pixel rectangle [x0,x1) x [y0,y1).
__global__ void k_clear_rect(double* film, int W, int x0, int y0, int x1, int y1) {
  const int rw = x1 - x0;
  const long long n = (long long)rw * (y1 - y0) * 3;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long p = i / 3;
  const int x = x0 + int(p % rw), y = y0 + int(p / rw);
  film[(size_t(y) * size_t(W) + size_t(x)) * 3 + size_t(i % 3)] = 0.0;
}

hipError_t launch_clear_rect(double* film, int W, int x0, int y0, int x1, int y1, hipStream_t st) {
  const long long n = (long long)(x1 - x0) * (y1 - y0) * 3;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_clear_rect, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, st, film, W, x0, y0, x1, y1);
  return hipGetLastError();
}

static inline uint32_t pad8(uint32_t n) { return (n + kNumXcd - 1) / kNumXcd * kNumXcd; }

template <typename R, bool C, int BLOCK, int WAVES = 0>
static hipError_t launch_render_t(const DevScene& s, const DevCamera& cam, const RenderArgs& a, hipStream_t st) {
  uint32_t waves = uint32_t(a.ntx) * uint32_t(a.nty) * uint32_t(a.nchunks);
  uint32_t nreal = (waves + (BLOCK / 64) - 1) / (BLOCK / 64);
  const dim3 grid(pad8(nreal)), blk(BLOCK);
  const bool surf = scene_has_surfaces(s);
  if constexpr (std::is_same<R, float>::value) {
    if (a.guide.cache.mix) {
      if (surf) hipLaunchKernelGGL((k_render<R, C, BLOCK, WAVES, true, true>), grid, blk, 0, st, s, cam, a, nreal);
      else hipLaunchKernelGGL((k_render<R, C, BLOCK, WAVES, false, true>), grid, blk, 0, st, s, cam, a, nreal);
      return hipGetLastError();
    }
  }
  if (surf) hipLaunchKernelGGL((k_render<R, C, BLOCK, WAVES, true, false>), grid, blk, 0, st, s, cam, a, nreal);
  else hipLaunchKernelGGL((k_render<R, C, BLOCK, WAVES, false, false>), grid, blk, 0, st, s, cam, a, nreal);
  return hipGetLastError();
}

hipError_t launch_render(const DevScene& s, const DevCamera& cam, const RenderArgs& a, int precision, bool count,
                         int block, int occupancy, hipStream_t st) {
  if (precision == 0 && !count && occupancy == 4) return launch_render_t<float, false, 128, 4>(s, cam, a, st);
  if (precision == 0 && !count && occupancy == 6) return launch_render_t<float, false, 128, 6>(s, cam, a, st);
#define NE_DISPATCH(BLK)                                                                       \
  if (precision == 1) {                                                                        \
    return count ? launch_render_t<double, true, BLK>(s, cam, a, st)                           \
                 : launch_render_t<double, false, BLK>(s, cam, a, st);                         \
  }                                                                                            \
  return count ? launch_render_t<float, true, BLK>(s, cam, a, st) : launch_render_t<float, false, BLK>(s, cam, a, st);
  if (block == 64) { NE_DISPATCH(64) }
  if (block == 256) { NE_DISPATCH(256) }
  NE_DISPATCH(128)
#undef NE_DISPATCH
}

hipError_t launch_reduce_partials(const RenderArgs& a, double* film, hipStream_t st) {
  size_t np = size_t(a.x1 - a.x0) * size_t(a.y1 - a.y0);
  hipLaunchKernelGGL(k_reduce_partials, dim3(uint32_t((np + 255) / 256)), dim3(256), 0, st, a, film);
  return hipGetLastError();
}

hipError_t launch_tonemap(const double* film, int npix, float inv_spp, float* pixels, hipStream_t st) {
  hipLaunchKernelGGL(k_tonemap, dim3(uint32_t((npix * 3 + 255) / 256)), dim3(256), 0, st, film, npix, inv_spp, pixels);
  return hipGetLastError();
}
hipError_t launch_tonemap_rect(const double* film, int W, int x0, int y0, int rw, int rh, float inv_spp, float* out,
                               hipStream_t st) {
  const int n = rw * rh * 3;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tonemap_rect, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, st, film, W, x0, y0, rw, rh,
                     inv_spp, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// majorant grid: `split` blocks per cell, max over the cell's trilinear
// support (the cell's voxels plus a one-voxel apron per side) times
// sigma_bar.  The blocks of a cell combine with an unsigned atomicMax, which
// orders non-negative floats like their values (maj is zeroed first).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_majorants(DevScene s, float* maj, int split) {
  const int M = s.mres;
  const int cell = int(blockIdx.x) / split, part = int(blockIdx.x) % split;
  const int c[3] = {cell % M, (cell / M) % M, cell / (M * M)};
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    int64_t r = s.res[a];
    lo[a] = int((int64_t(c[a]) * r) / M) - 1;
    lo[a] = lo[a] < 0 ? 0 : lo[a];
    int h = int((int64_t(c[a] + 1) * r + M - 1) / M) + 1;
    hi[a] = h < int(r) - 1 ? h : int(r) - 1;
  }
  const int64_t nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
  const int64_t total = nx * ny * nz;
  float m = 0.0f;
  for (int64_t i = int64_t(part) * blockDim.x + threadIdx.x; i < total; i += int64_t(split) * blockDim.x) {
    const int64_t x = lo[0] + i % nx, y = lo[1] + (i / nx) % ny, z = lo[2] + i / (nx * ny);
    m = xmax(m, grid_at(s, int(x), int(y), int(z)));
  }
  __shared__ float red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] = xmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(reinterpret_cast<unsigned int*>(maj + cell), __float_as_uint(red[0] * s.sigma_bar));
}

// Row-major (x-fastest) density -> bricked layout, zeros in the padding: one
// thread per padded voxel, in storage order.
__global__ void k_brick_grid(const float* lin, float* out, int rx, int ry, int rz, uint32_t px, uint32_t py,
                             uint64_t total) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const uint32_t inner = uint32_t(i & (kBrickFloats - 1u));
  const uint64_t b = i / kBrickFloats;
  const uint32_t nbx = px / 4u, nby = py / 4u;
  const uint32_t x = uint32_t(b % nbx) * 4u + (inner & 3u);
  const uint32_t y = uint32_t((b / nbx) % nby) * 4u + ((inner >> 2) & 3u);
  const uint32_t z = uint32_t(b / (uint64_t(nbx) * nby)) * 2u + (inner >> 4);
  out[i] = (x < uint32_t(rx) && y < uint32_t(ry) && z < uint32_t(rz))
               ? lin[(uint64_t(z) * uint32_t(ry) + y) * uint32_t(rx) + x]
               : 0.0f;
}

hipError_t launch_brick_grid(const float* lin, float* out, const int res[3], hipStream_t st) {
  const uint32_t px = grid_pad_x(res[0]), py = grid_pad_y(res[1]), pz = grid_pad_z(res[2]);
  const uint64_t total = uint64_t(px) * py * pz;
  hipLaunchKernelGGL(k_brick_grid, dim3(uint32_t((total + 255) / 256)), dim3(256), 0, st, lin, out, res[0], res[1],
                     res[2], px, py, total);
  return hipGetLastError();
}

hipError_t launch_majorants(const DevScene& s, float* maj, hipStream_t st) {
  const uint32_t cells = uint32_t(s.mres) * uint32_t(s.mres) * uint32_t(s.mres);
  const int split = int(cells >= 2048u ? 1u : 2048u / cells);  // at least ~2048 blocks
  hipError_t e = hipMemsetAsync(maj, 0, cells * sizeof(float), st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_majorants, dim3(cells * uint32_t(split)), dim3(256), 0, st, s, maj, split);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// parity kernels
// ---------------------------------------------------------------------------
__global__ void k_test_primitive(int kind, int n, const float* rays, const float* prims, int32_t* hit, float* t0,
                                 float* t1) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Vec3<float> o = v3f<float>(rays + 6 * i), d = v3f<float>(rays + 6 * i + 3);
  float a = __builtin_huge_valf(), b = -__builtin_huge_valf();
  bool h = false;
  if (kind == 0) {
    h = ref_aabb(v3f<float>(prims + 6 * i), v3f<float>(prims + 6 * i + 3), o, d, a, b);
  } else if (kind == 1) {
    const float* q = prims + 9 * i;
    h = ref_triangle(v3f<float>(q), v3f<float>(q + 3), v3f<float>(q + 6), o, d, a);
    b = a;
  } else if (kind == 2) {
    h = ref_sphere(v3f<float>(prims + 4 * i), prims[4 * i + 3], o, d, a, b);
  } else {
    const float* q = prims + 9 * i;
    h = ref_rect(v3f<float>(q), v3f<float>(q + 3), v3f<float>(q + 6), o, d, a);
    b = a;
  }
  hit[i] = h ? 1 : 0;
  t0[i] = a;
  t1[i] = b;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_test_closest(DevScene s, int n, const float* rays, float tmin, float tmax,
                                                         float* t, int32_t* type, int32_t* index) {
  __shared__ uint32_t lds_stack[kShortStack * BLOCK];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Count<false> cnt;
  SurfHit<float> h;
  bool ok = traverse<float, false, false>(s, v3f<float>(rays + 6 * i), v3f<float>(rays + 6 * i + 3), tmin, tmax,
                                          lds_stack + threadIdx.x, BLOCK, h, cnt);
  t[i] = ok ? h.t : __builtin_huge_valf();
  type[i] = ok ? h.type : -1;
  index[i] = ok ? int32_t(h.key) : -1;
}

template <typename R, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_test_transmittance(DevScene s, int n, const float* rays, const float* tmax,
                                                               uint64_t seed, double* out) {
  __shared__ uint32_t lds_stack[kShortStack * BLOCK];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Count<false> cnt;
  PathRng rng;
  rng.init(seed, uint32_t(i), 0, 0);
  R tr = transmittance<R, false>(s, v3f<R>(rays + 6 * i), v3f<R>(rays + 6 * i + 3), R(tmax[i]), rng,
                                 lds_stack + threadIdx.x, BLOCK, cnt);
  out[i] = double(tr);
}

template <typename R>
__global__ void k_test_delta(DevScene s, int n, const float* rays, uint64_t seed, int32_t* scattered, double* tout) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Count<false> cnt;
  PathRng rng;
  rng.init(seed, uint32_t(i), 0, 0);
  Vec3<R> o = v3f<R>(rays + 6 * i), d = v3f<R>(rays + 6 * i + 3);
  Vec3<R> oo = xf_point<R>(s.inv, o), dd = xf_vec<R>(s.inv, d);
  R ta, tb, th;
  scattered[i] = 0;
  tout[i] = 0;
  if (box_interval<R>(s, oo, dd, ta, tb)) {
    ta = xmax(ta, R(0));
    if (ta < tb && delta_track<R, false>(s, o, d, ta, tb, rng, cnt, th)) { scattered[i] = 1; tout[i] = double(th); }
  }
}

// Even points take the 32-bit-offset lookup when the grid allows it, odd
// points the general one, so one call checks both against the caller's values.
__global__ void k_test_density(DevScene s, int n, const float* gcs, float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Vec3<float> p = v3f<float>(gcs + 3 * i);
  out[i] = (s.grid_small && (i & 1) == 0) ? density_gcs<true>(s, p) : density_gcs<false>(s, p);
}

__global__ void k_test_bsdf(int kind, int n, const float* in0, const float* in1, const float* in2, const float* in3,
                            float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (kind == 0) {  // HG eval: a, b, g
    out[i] = ref_hg_eval(in2[i], v3f<float>(in0 + 3 * i), v3f<float>(in1 + 3 * i));
  } else if (kind == 1) {  // HG sample local: u, g
    Vec3<float> w = ref_hg_sample(in1[i], in0[2 * i], in0[2 * i + 1]);
    out[3 * i] = w.x; out[3 * i + 1] = w.y; out[3 * i + 2] = w.z;
  } else if (kind == 2) {  // GGX D: h, alpha
    out[i] = ref_ggx_D(in1[i], v3f<float>(in0 + 3 * i));
  } else if (kind == 3) {  // GGX pdf: wo, h, alpha
    out[i] = ref_ggx_pdf(in2[i], v3f<float>(in0 + 3 * i), v3f<float>(in1 + 3 * i));
  } else if (kind == 4) {  // glossy: mat(5), wo, wi, u(3)
    Glossy<float> m;
    m.albedo = v3f<float>(in0 + 5 * i);
    m.alpha = in0[5 * i + 3] * in0[5 * i + 3];
    m.metallic = in0[5 * i + 4];
    Vec3<float> wo = v3f<float>(in1 + 3 * i), wi = v3f<float>(in2 + 3 * i);
    Vec3<float> f = m.eval(wo, wi);
    float p = m.pdf(wo, wi);
    Vec3<float> sm = m.sample(wo, in3[3 * i], in3[3 * i + 1], in3[3 * i + 2]);
    float* o = out + 7 * i;
    o[0] = f.x; o[1] = f.y; o[2] = f.z; o[3] = p; o[4] = sm.x; o[5] = sm.y; o[6] = sm.z;
  } else if (kind == 5) {  // forward HG: d, u, g
    float g = in2[i];
    Vec3<float> d = v3f<float>(in0 + 3 * i);
    Vec3<float> w = hg_sample<float>(g, d, in1[2 * i], in1[2 * i + 1]);
    float* o = out + 4 * i;
    o[0] = w.x; o[1] = w.y; o[2] = w.z; o[3] = hg_p<float>(g, dot(d, w));
  }
}

__global__ void k_test_rng(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t job, int n, float* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  PathRng r;
  r.init(seed, pixel, sample, job);
  for (int i = 0; i < n; ++i) out[i] = r.next();
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_li(DevScene s, int n, const float* rays, int max_bounces, uint64_t seed,
                                               uint32_t sample, float* out) {
  __shared__ uint32_t lds_stack[kShortStack * BLOCK];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Count<false> cnt;
  PathRng rng;
  rng.init(seed, uint32_t(i), sample, 0);
  Vec3<float> L = trace_path<float, false>(s, v3f<float>(rays + 6 * i), v3f<float>(rays + 6 * i + 3), max_bounces,
                                           rng, lds_stack + threadIdx.x, BLOCK, cnt);
  out[3 * i] = L.x; out[3 * i + 1] = L.y; out[3 * i + 2] = L.z;
}

static inline dim3 grid1(int n, int b) { return dim3(uint32_t((n + b - 1) / b)); }

hipError_t launch_test_primitive(int kind, int n, const float* rays, const float* prims, int32_t* hit, float* t0,
                                 float* t1, hipStream_t st) {
  hipLaunchKernelGGL(k_test_primitive, grid1(n, 256), dim3(256), 0, st, kind, n, rays, prims, hit, t0, t1);
  return hipGetLastError();
}
hipError_t launch_test_closest(const DevScene& s, int n, const float* rays, float tmin, float tmax, float* t,
                               int32_t* type, int32_t* index, hipStream_t st) {
  hipLaunchKernelGGL(k_test_closest<128>, grid1(n, 128), dim3(128), 0, st, s, n, rays, tmin, tmax, t, type, index);
  return hipGetLastError();
}
hipError_t launch_test_transmittance(const DevScene& s, int n, const float* rays, const float* tmax, uint64_t seed,
                                     int precision, double* out, hipStream_t st) {
  if (precision == 1)
    hipLaunchKernelGGL((k_test_transmittance<double, 128>), grid1(n, 128), dim3(128), 0, st, s, n, rays, tmax, seed,
                       out);
  else
    hipLaunchKernelGGL((k_test_transmittance<float, 128>), grid1(n, 128), dim3(128), 0, st, s, n, rays, tmax, seed,
                       out);
  return hipGetLastError();
}
hipError_t launch_test_delta(const DevScene& s, int n, const float* rays, uint64_t seed, int precision,
                             int32_t* scattered, double* t, hipStream_t st) {
  if (precision == 1)
    hipLaunchKernelGGL(k_test_delta<double>, grid1(n, 128), dim3(128), 0, st, s, n, rays, seed, scattered, t);
  else
    hipLaunchKernelGGL(k_test_delta<float>, grid1(n, 128), dim3(128), 0, st, s, n, rays, seed, scattered, t);
  return hipGetLastError();
}
hipError_t launch_test_density(const DevScene& s, int n, const float* gcs, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_test_density, grid1(n, 256), dim3(256), 0, st, s, n, gcs, out);
  return hipGetLastError();
}
hipError_t launch_test_bsdf(int kind, int n, const float* in0, const float* in1, const float* in2, const float* in3,
                            float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_test_bsdf, grid1(n, 256), dim3(256), 0, st, kind, n, in0, in1, in2, in3, out);
  return hipGetLastError();
}
hipError_t launch_test_rng(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t job, int n, float* out,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_test_rng, dim3(1), dim3(64), 0, st, seed, pixel, sample, job, n, out);
  return hipGetLastError();
}
hipError_t launch_li(const DevScene& s, int n, const float* rays, int max_bounces, uint64_t seed, uint32_t sample,
                     float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_li<128>, grid1(n, 128), dim3(128), 0, st, s, n, rays, max_bounces, seed, sample, out);
  return hipGetLastError();
}

}  // namespace ne
