// api.cpp -- implementation of the C ABI in include/ne_gpu.h: contexts,
// scene flattening and upload (SAH BVH + majorant grid), films, render
// dispatch, counters and the parity entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <chrono>
#include <thread>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ne_gpu.h"
#include "bvh_build.h"
#include "ne_launch.h"
#include "ne_math.h"
#include "ne_scene.h"

using namespace ne;

struct ne_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::string err;
  DevCounters* d_counters = nullptr;
  DevCounters h_counters{};
  bool counters_valid = false;
  double* d_partial = nullptr;
  size_t partial_bytes = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  // One reference for the caller's handle plus one per live scene/film: the
  // context outlives every object created on it, whatever the release order.
  std::atomic<int> refs{1};
  void* wf_buf = nullptr;  // wavefront path pool (grown on demand)
  uint32_t wf_P = 0;
  uint32_t* h_status = nullptr;  // pinned status word read by the wavefront loop
  std::vector<hipEvent_t> ev_pool;  // per-launch timing events (NE_FLAG_STAGE_TIME)
  std::vector<int8_t> ev_stages;    // stage of the launch after each recorded event
  bool stage_timed = false;
  uint32_t refill_blocks = 0;  // resident lane-refill blocks on the device (CUs x blocks per CU)
  void* wfg_buf = nullptr;      // guided-mode training records (grown on demand)
  uint32_t wfg_P = 0;
  void* sbuf = nullptr;         // wavefront per-sample radiance (grown on demand)
  size_t sbuf_bytes = 0;
  double* scratch = nullptr;    // this device's copy of the static scratch array (ne_render_noalloc)
};

struct ne_scene {
  ne_ctx* ctx = nullptr;
  DevScene dev{};
  std::vector<void*> allocs;
  uint64_t bytes = 0;
  ne_scene_info info{};
};

struct ne_film {
  ne_ctx* ctx = nullptr;
  int W = 0, H = 0;
  double* d_sum = nullptr;
  long long* d_fixed = nullptr;  // wavefront fixed-point accumulator, flushed into d_sum per render
};

struct ne_guide {
  ne_ctx* ctx = nullptr;
  ne_guide_config cfg{};
  GuideNet net{};
  GuideBatch batch{};
  GuideTrain train{};
  std::vector<void*> allocs;
  size_t table_n = 0;
  uint64_t steps = 0, samples = 0;
  float last_loss = 0.0f;
  GuideCache cache{};        // tabulated proposal (cache.mix == nullptr: none)
  float* cache_mix = nullptr;
  uint64_t version = 1;      // bumped by every parameter change
  uint64_t cache_version = 0;  // parameters the tabulation was made from
};

// the guide's proposal tabulation and its use in a render (defined with the guide API)
static ne_status guide_refresh_cache(ne_ctx* c, ne_guide* g);
static GuideUse guide_use(const ne_guide* g);

// ---- static device scratch ------------------------------------------------------
// The allocation-free render (ne_render_noalloc) keeps its megakernel's
// per-chunk partial sums in this fixed array of the code object: each device
// holds its own copy until the process ends.  In host code `scratch` names a
// host-side stand-in, never a device address: kernels, copies and fills take
// the per-device address that hipGetSymbolAddress returns (resolved once per
// device, with that device current, when a context is created on it).
constexpr size_t kStaticScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

thread_local std::string g_err;

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_dev;  // per device index; nullptr until resolved

// The calling thread's current device must be `device`.
hipError_t resolve_scratch(int device, double** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (size_t(device) >= g_scratch_dev.size()) g_scratch_dev.resize(size_t(device) + 1, nullptr);
  if (!g_scratch_dev[size_t(device)]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[size_t(device)] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[size_t(device)];
  return hipSuccess;
}

ne_status fail(ne_ctx* c, ne_status st, const std::string& msg) {
  if (c) c->err = msg; else g_err = msg;
  return st;
}

#define NE_HIP(ctx, call)                                                                          \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return fail(ctx, e_ == hipErrorOutOfMemory ? NE_ERR_OUT_OF_MEMORY : NE_ERR_HIP,              \
                  std::string(#call) + ": " + hipGetErrorString(e_));                             \
  } while (0)

template <class T>
ne_status dev_upload(ne_ctx* c, ne_scene* s, const T* src, size_t count, const T** dst) {
  *dst = nullptr;
  if (count == 0) return NE_OK;
  void* p = nullptr;
  size_t bytes = count * sizeof(T);
  NE_HIP(c, hipMalloc(&p, bytes));
  s->allocs.push_back(p);
  s->bytes += bytes;
  NE_HIP(c, hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, c->stream));
  *dst = static_cast<const T*>(p);
  return NE_OK;
}

// scoped device buffers for the parity entry points
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
};

template <class T>
ne_status tmp_in(ne_ctx* c, DevBuf& b, const T* src, size_t count) {
  if (count == 0 || src == nullptr) return NE_OK;
  NE_HIP(c, hipMalloc(&b.p, count * sizeof(T)));
  NE_HIP(c, hipMemcpyAsync(b.p, src, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return NE_OK;
}
template <class T>
ne_status tmp_out(ne_ctx* c, DevBuf& b, size_t count) {
  NE_HIP(c, hipMalloc(&b.p, std::max<size_t>(count, 1) * sizeof(T)));
  return NE_OK;
}
template <class T>
ne_status tmp_back(ne_ctx* c, T* dst, const DevBuf& b, size_t count) {
  NE_HIP(c, hipMemcpyAsync(dst, b.p, count * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  return NE_OK;
}

inline float len3(const float* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// Same operation order as the device/oracle: length(cross(eu, ev)).
inline float rect_area(const float* eu, const float* ev) {
  float c[3] = {eu[1] * ev[2] - ev[1] * eu[2], eu[2] * ev[0] - ev[2] * eu[0], eu[0] * ev[1] - ev[0] * eu[1]};
  return std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
}

// Empty-space skipping over an M^3 majorant grid: every empty cell (majorant
// 0) gets -D, D the L1 distance to the nearest non-empty cell (3M when there
// is none).  The L1 distance transform is separable: per axis, a forward and
// a backward sweep of d[k] = min(d[k], d[k -/+ 1] + 1).
void encode_empty_distance(std::vector<float>& maj, int M) {
  const int far = 3 * M;
  std::vector<int> d(maj.size());
  for (size_t i = 0; i < maj.size(); ++i) d[i] = maj[i] > 0.0f ? 0 : far;
  const size_t stride[3] = {1, size_t(M), size_t(M) * size_t(M)};
  for (int ax = 0; ax < 3; ++ax) {
    const size_t st = stride[ax], su = stride[(ax + 1) % 3], sv = stride[(ax + 2) % 3];
    for (int u = 0; u < M; ++u)
      for (int v = 0; v < M; ++v) {
        int* line = d.data() + size_t(u) * su + size_t(v) * sv;
        for (int k = 1; k < M; ++k) line[k * st] = std::min(line[k * st], line[(k - 1) * st] + 1);
        for (int k = M - 2; k >= 0; --k) line[k * st] = std::min(line[k * st], line[(k + 1) * st] + 1);
      }
  }
  for (size_t i = 0; i < maj.size(); ++i)
    if (!(maj[i] > 0.0f)) maj[i] = -float(std::min(d[i], far));
}

// Sampling tables of an environment map: func = luminance * sin(theta of the
// row centre); per-row CDFs over the row's W texels and the marginal CDF over
// rows, summed in FP64 and stored in FP32 with exact 0 and 1 ends (uniform
// where a row or the whole map is black).  inv_int = W H / sum(func).
struct EnvTables {
  std::vector<float> func, cdf, marg;
  float inv_int = 0.0f;
};
void build_env_tables(const float* rgb, int W, int H, EnvTables& t) {
  t.func.assign(size_t(W) * H, 0.0f);
  t.cdf.assign(size_t(H) * (W + 1), 0.0f);
  t.marg.assign(size_t(H) + 1, 0.0f);
  std::vector<double> row(static_cast<size_t>(H)), pre(static_cast<size_t>(W) + 1);
  double total = 0.0;
  for (int j = 0; j < H; ++j) {
    const double st = std::sin(kPI * (double(j) + 0.5) / double(H));
    double acc = 0.0;
    pre[0] = 0.0;
    for (int i = 0; i < W; ++i) {
      const float* p = rgb + 3 * (size_t(j) * W + i);
      const float f = float((0.2126 * p[0] + 0.7152 * p[1] + 0.0722 * p[2]) * st);
      t.func[size_t(j) * W + i] = f;
      acc += double(f);
      pre[size_t(i) + 1] = acc;
    }
    float* c = &t.cdf[size_t(j) * (W + 1)];
    for (int i = 1; i < W; ++i) c[i] = acc > 0 ? float(pre[size_t(i)] / acc) : float(double(i) / W);
    c[0] = 0.0f;
    c[W] = 1.0f;
    row[size_t(j)] = acc;
    total += acc;
  }
  double acc = 0.0;
  for (int j = 1; j < H; ++j) {
    acc += row[size_t(j) - 1];
    t.marg[size_t(j)] = total > 0 ? float(acc / total) : float(double(j) / H);
  }
  t.marg[0] = 0.0f;
  t.marg[size_t(H)] = 1.0f;
  t.inv_int = total > 0 ? float(double(W) * double(H) / total) : 0.0f;
}

ne_status validate_desc(ne_ctx* c, const ne_scene_desc* d) {
  if (!d) return fail(c, NE_ERR_INVALID_ARG, "scene descriptor is NULL");
  if (d->n_tri && !d->tri_v) return fail(c, NE_ERR_INVALID_ARG, "tri_v is NULL");
  if (d->n_sph && (!d->sph || !d->sph_mat)) return fail(c, NE_ERR_INVALID_ARG, "sphere arrays are NULL");
  if (d->n_rect && (!d->rect || !d->rect_mat)) return fail(c, NE_ERR_INVALID_ARG, "rectangle arrays are NULL");
  if ((d->n_tri || d->n_sph || d->n_rect) && (d->n_mat == 0 || !d->mat))
    return fail(c, NE_ERR_INVALID_ARG, "surfaces need at least one material");
  if (d->n_tri > kItemFirstMask || d->n_sph > kItemFirstMask || d->n_rect > kItemFirstMask)
    return fail(c, NE_ERR_INVALID_ARG, "too many primitives of one type (limit 2^26)");
  auto chk_mat = [&](int32_t m) { return m >= 0 && uint32_t(m) < d->n_mat; };
  // Every coordinate must be finite: the BVH builder cannot bound a primitive
  // with an infinite or NaN extent, and a box that misses it would drop hits.
  auto finite = [](const float* v, size_t n) {
    for (size_t k = 0; k < n; ++k)
      if (!std::isfinite(v[k])) return false;
    return true;
  };
  for (uint32_t i = 0; i < d->n_tri; ++i) {
    if (!chk_mat(d->tri_mat ? d->tri_mat[i] : 0)) return fail(c, NE_ERR_INVALID_ARG, "triangle material out of range");
    if (!finite(d->tri_v + 9 * size_t(i), 9)) return fail(c, NE_ERR_INVALID_ARG, "triangle vertices must be finite");
  }
  for (uint32_t i = 0; i < d->n_sph; ++i) {
    if (!chk_mat(d->sph_mat[i])) return fail(c, NE_ERR_INVALID_ARG, "sphere material out of range");
    if (!finite(d->sph + 4 * size_t(i), 4)) return fail(c, NE_ERR_INVALID_ARG, "sphere centre and radius must be finite");
    if (!(d->sph[4 * i + 3] > 0)) return fail(c, NE_ERR_INVALID_ARG, "sphere radius must be > 0");
  }
  for (uint32_t i = 0; i < d->n_rect; ++i) {
    if (!chk_mat(d->rect_mat[i])) return fail(c, NE_ERR_INVALID_ARG, "rectangle material out of range");
    if (!finite(d->rect + 12 * size_t(i), 12)) return fail(c, NE_ERR_INVALID_ARG, "rectangle corners must be finite");
    if (!(rect_area(d->rect + 12 * i + 3, d->rect + 12 * i + 6) > 0))
      return fail(c, NE_ERR_INVALID_ARG, "degenerate rectangle");
  }
  if (d->n_point && (!d->point || !d->point_mat)) return fail(c, NE_ERR_INVALID_ARG, "point arrays are NULL");
  if (d->n_point > kItemFirstMask) return fail(c, NE_ERR_INVALID_ARG, "too many point lights (limit 2^26)");
  for (uint32_t i = 0; i < d->n_point; ++i) {
    if (!chk_mat(d->point_mat[i]) || d->mat[12 * size_t(d->point_mat[i])] != float(NE_MAT_EMITTER))
      return fail(c, NE_ERR_INVALID_ARG, "a point light needs an emitter material");
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(d->point[3 * size_t(i) + a])) return fail(c, NE_ERR_INVALID_ARG, "point light position");
  }
  for (uint32_t i = 0; i < d->n_mat; ++i) {
    float t = d->mat[12 * i];
    if (t != float(NE_MAT_GLOSSY) && t != float(NE_MAT_EMITTER))
      return fail(c, NE_ERR_INVALID_ARG, "unknown material type");
  }
  if (d->has_volume) {
    const bool homog = d->vol_homogeneous != 0;
    if (!homog && !d->grid) return fail(c, NE_ERR_INVALID_ARG, "volume grid is NULL");
    for (int a = 0; a < 3 && !homog; ++a)
      if (d->vol_res[a] <= 0 || d->vol_res[a] > 4096) return fail(c, NE_ERR_INVALID_ARG, "volume resolution out of range");
    if (!homog && (d->maj_res < 1 || d->maj_res > 256))
      return fail(c, NE_ERR_INVALID_ARG, "majorant resolution out of range");
    float st = 0;
    for (int a = 0; a < 3; ++a) st += d->vol_sigma_s[a] + d->vol_sigma_a[a];
    if (!(st > 0) || !(d->vol_mult > 0)) return fail(c, NE_ERR_INVALID_ARG, "volume extinction must be > 0");
    for (int a = 0; a < 3; ++a)
      if (!(d->vol_sigma_s[a] + d->vol_sigma_a[a] > 0)) return fail(c, NE_ERR_INVALID_ARG, "per-channel extinction must be > 0");
    if (!(d->vol_g > -1.0f && d->vol_g < 1.0f)) return fail(c, NE_ERR_INVALID_ARG, "HG g must be in (-1, 1)");
    if (!std::isfinite(d->vol_mult)) return fail(c, NE_ERR_INVALID_ARG, "density multiplier must be finite");
    // Tracking loops step by -log(1 - xi) / majorant: a NaN or infinite voxel
    // would stall them, and negative densities break the majorant bound.
    const size_t nvox = homog ? 0 : size_t(d->vol_res[0]) * d->vol_res[1] * d->vol_res[2];
    for (size_t i = 0; i < nvox; ++i)
      if (!(d->grid[i] >= 0.0f) || !std::isfinite(d->grid[i]))
        return fail(c, NE_ERR_INVALID_ARG, "grid densities must be finite and >= 0");
  }
  if (d->env_type < 0 || d->env_type > 3) return fail(c, NE_ERR_INVALID_ARG, "unknown environment type");
  if (d->env_type == NE_ENV_MAP) {
    if (!d->env_map) return fail(c, NE_ERR_INVALID_ARG, "environment map is NULL");
    if (d->env_w < 1 || d->env_h < 1 || d->env_w > 16384 || d->env_h > 16384)
      return fail(c, NE_ERR_INVALID_ARG, "environment map size out of [1, 16384]");
    const size_t n = size_t(d->env_w) * size_t(d->env_h) * 3;
    for (size_t i = 0; i < n; ++i)
      if (!(d->env_map[i] >= 0.0f) || !std::isfinite(d->env_map[i]))
        return fail(c, NE_ERR_INVALID_ARG, "environment map values must be finite and >= 0");
  }
  if (d->has_sun && !(len3(d->sun_dir) > 0)) return fail(c, NE_ERR_INVALID_ARG, "sun direction is zero");
  return NE_OK;
}

ne_status check_params(ne_ctx* c, const ne_render_params* p) {
  if (!p) return fail(c, NE_ERR_INVALID_ARG, "params is NULL");
  if (p->width <= 0 || p->height <= 0 || p->width > 32768 || p->height > 32768)
    return fail(c, NE_ERR_INVALID_ARG, "image size out of range");
  if (p->x0 < 0 || p->y0 < 0 || p->x1 > p->width || p->y1 > p->height || p->x0 >= p->x1 || p->y0 >= p->y1)
    return fail(c, NE_ERR_INVALID_ARG, "pixel rectangle out of range or empty");
  if (p->spp <= 0) return fail(c, NE_ERR_INVALID_ARG, "spp must be > 0");
  if (p->max_bounces < 0 || p->max_bounces > kMaxVertices)
    return fail(c, NE_ERR_INVALID_ARG, "max_bounces out of range [0, 2047]");
  if (p->precision != NE_FP32 && p->precision != NE_FP64) return fail(c, NE_ERR_INVALID_ARG, "unknown precision");
  if (p->block_size != 0 && p->block_size != 64 && p->block_size != 128 && p->block_size != 256)
    return fail(c, NE_ERR_INVALID_ARG, "block_size must be 0, 64, 128 or 256");
  if (p->samples_per_thread < 0) return fail(c, NE_ERR_INVALID_ARG, "samples_per_thread must be >= 0");
  if (p->occupancy != 0 && p->occupancy != 4 && p->occupancy != 6)
    return fail(c, NE_ERR_INVALID_ARG, "occupancy must be 0, 4 or 6");
  if (p->mode < NE_MODE_AUTO || p->mode > NE_MODE_WAVEFRONT) return fail(c, NE_ERR_INVALID_ARG, "unknown mode");
  if (p->mode == NE_MODE_WAVEFRONT && p->precision != NE_FP32)
    return fail(c, NE_ERR_INVALID_ARG, "the wavefront pipeline is FP32 only");
  if (p->pool_size < 0 || p->pool_size > (1 << 28))
    return fail(c, NE_ERR_INVALID_ARG, "pool_size must be in [0, 2^28]");
  return NE_OK;
}

// World-space bounds of every surface primitive: triangles, then spheres,
// then rectangles (the BVH builder's input index space).
std::vector<PrimInput> prim_bounds(const ne_scene_desc* d) {
  const uint32_t nt = d->n_tri, ns = d->n_sph, nr = d->n_rect;
  std::vector<PrimInput> prims(size_t(nt) + ns + nr);
  for (uint32_t i = 0; i < nt; ++i) {
    const float* v = d->tri_v + 9 * size_t(i);
    for (int a = 0; a < 3; ++a) {
      prims[i].bmin[a] = std::min(v[a], std::min(v[3 + a], v[6 + a]));
      prims[i].bmax[a] = std::max(v[a], std::max(v[3 + a], v[6 + a]));
    }
    prims[i].type = PRIM_TRI;
  }
  for (uint32_t i = 0; i < ns; ++i) {
    const float* q = d->sph + 4 * size_t(i);
    for (int a = 0; a < 3; ++a) { prims[nt + i].bmin[a] = q[a] - q[3]; prims[nt + i].bmax[a] = q[a] + q[3]; }
    prims[nt + i].type = PRIM_SPHERE;
  }
  for (uint32_t i = 0; i < nr; ++i) {
    const float* r = d->rect + 12 * size_t(i);
    for (int a = 0; a < 3; ++a) {
      float c0 = r[a], c1 = r[a] + r[3 + a], c2 = r[a] + r[6 + a], c3 = r[a] + r[3 + a] + r[6 + a];
      prims[nt + ns + i].bmin[a] = std::min(std::min(c0, c1), std::min(c2, c3));
      prims[nt + ns + i].bmax[a] = std::max(std::max(c0, c1), std::max(c2, c3));
    }
    prims[nt + ns + i].type = PRIM_RECT;
  }
  return prims;
}

DevCamera make_camera(const ne_camera_desc* cd) {
  // Camera::Camera (src/core/Camera.cpp:7-26) with focus distance 1 and no lens.
  auto sub = [](const float* a, const float* b, float* o) { for (int i = 0; i < 3; ++i) o[i] = a[i] - b[i]; };
  auto nrm = [](float* v) {
    float d = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    float s = 1.0f / std::sqrt(d);
    v[0] *= s; v[1] *= s; v[2] *= s;
  };
  float front[3], side[3];
  sub(cd->position, cd->look_at, front);
  nrm(front);
  const float* up = cd->up;
  side[0] = up[1] * front[2] - front[1] * up[2];
  side[1] = up[2] * front[0] - front[2] * up[0];
  side[2] = up[0] * front[1] - front[0] * up[1];
  nrm(side);
  const float focus = 1.0f;
  float halfHeight = std::tan(float(cd->vfov * kPI / 180.0) / 2.0f);
  float halfWidth = cd->aspect * halfHeight;
  DevCamera c;
  for (int i = 0; i < 3; ++i) {
    c.pos[i] = cd->position[i];
    float a = focus * halfWidth * (-side[i]);
    float b = focus * halfHeight * up[i];
    float f = focus * (-front[i]);
    c.lower_left[i] = ((cd->position[i] - a) - b) - f;
    c.horizontal[i] = 2.0f * focus * halfWidth * (-side[i]);
    c.vertical[i] = 2.0f * focus * halfHeight * up[i];
  }
  return c;
}

}  // namespace

namespace ne {
// Error of a context-free entry point (scene files), read by ne_last_error(NULL).
ne_status host_fail(ne_status st, const std::string& msg) { return fail(nullptr, st, msg); }
}

extern "C" {

int32_t ne_sizeof_scene_desc(void) { return int32_t(sizeof(ne_scene_desc)); }

const char* ne_last_error(const ne_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

ne_status ne_bvh_build(const ne_scene_desc* d, float* nodes, uint32_t* refs, uint32_t* n_nodes, uint32_t* n_refs,
                       uint32_t* depth) {
  ne_status st = validate_desc(nullptr, d);
  if (st != NE_OK) return st;
  BvhOutput bvh;
  build_bvh(prim_bounds(d), bvh);
  if (bvh.quant_failed) return fail(nullptr, NE_ERR_INVALID_ARG, "BVH quantization failed: non-finite bounds");
  if (n_nodes) *n_nodes = uint32_t(bvh.nodes.size() / kNodeFloats);
  if (n_refs) *n_refs = uint32_t(bvh.refs.size());
  if (depth) *depth = bvh.max_depth;
  if (nodes) std::memcpy(nodes, bvh.nodes.data(), bvh.nodes.size() * sizeof(float));
  if (refs) {
    const uint32_t nt = d->n_tri, ns = d->n_sph;
    for (size_t i = 0; i < bvh.refs.size(); ++i) {
      uint32_t p = bvh.refs[i];
      refs[i] = p < nt ? p : (p < nt + ns ? ((uint32_t(PRIM_SPHERE) << 30) | (p - nt))
                                          : ((uint32_t(PRIM_RECT) << 30) | (p - nt - ns)));
    }
  }
  return NE_OK;
}

static void ctx_free(ne_ctx* c);

ne_status ne_ctx_create(int32_t device, ne_ctx** out) {
  if (!out) return fail(nullptr, NE_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(nullptr, NE_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= n) return fail(nullptr, NE_ERR_INVALID_ARG, "device index out of range");
  ne_ctx* c = new ne_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->d_counters, sizeof(DevCounters)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_status, 64 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess ||
      hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return fail(nullptr, NE_ERR_HIP, "context setup failed");
  }
  if (resolve_scratch(device, &c->scratch) != hipSuccess) {
    ctx_free(c);
    return fail(nullptr, NE_ERR_HIP, "static scratch address lookup failed");
  }
  *out = c;
  return NE_OK;
}

static void ctx_free(ne_ctx* c) {
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->d_partial) (void)hipFree(c->d_partial);
  if (c->d_counters) (void)hipFree(c->d_counters);
  if (c->wf_buf) (void)hipFree(c->wf_buf);
  if (c->h_status) (void)hipHostFree(c->h_status);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->wfg_buf) (void)hipFree(c->wfg_buf);
  if (c->sbuf) (void)hipFree(c->sbuf);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

static void ctx_unref(ne_ctx* c) {
  if (c->refs.fetch_sub(1) == 1) ctx_free(c);
}

void ne_ctx_destroy(ne_ctx* c) {
  if (!c) return;
  ctx_unref(c);
}

ne_status ne_synchronize(ne_ctx* c) {
  if (!c) return fail(nullptr, NE_ERR_INVALID_ARG, "ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

ne_status ne_scene_upload(ne_ctx* c, const ne_scene_desc* d, ne_scene** out) {
  if (!c || !out) return fail(c, NE_ERR_INVALID_ARG, "ctx/out is NULL");
  *out = nullptr;
  std::lock_guard<std::mutex> lk(c->mu);
  ne_status st = validate_desc(c, d);
  if (st != NE_OK) return st;
  NE_HIP(c, hipSetDevice(c->device));
  ne_scene* s = new ne_scene();
  s->ctx = c;
  auto bail = [&](ne_status e) {
    (void)hipStreamSynchronize(c->stream);
    for (void* p : s->allocs) (void)hipFree(p);
    delete s;
    return e;
  };
  // a failing HIP call releases the partly built scene
#define NE_HIP_BAIL(call)                                                                          \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return bail(fail(c, e_ == hipErrorOutOfMemory ? NE_ERR_OUT_OF_MEMORY : NE_ERR_HIP,           \
                       std::string(#call) + ": " + hipGetErrorString(e_)));                       \
  } while (0)
  DevScene& ds = s->dev;
  // ---- surfaces: BVH over triangles, spheres, rectangles.  Each type is
  // stored in leaf order, and leaf children point at their type's storage.
  const uint32_t nt = d->n_tri, ns = d->n_sph, nr = d->n_rect;
  BvhOutput bvh;
  build_bvh(prim_bounds(d), bvh);
  if (bvh.quant_failed) return bail(fail(c, NE_ERR_INVALID_ARG, "BVH quantization failed: non-finite bounds"));
  if (bvh.max_stack > uint32_t(kShortStack + kSpillStack)) return bail(fail(c, NE_ERR_UNSUPPORTED, "BVH too deep"));
  if (bvh.max_leaf > kItemMaxLeaf)
    return bail(fail(c, NE_ERR_UNSUPPORTED, "degenerate geometry: a BVH leaf holds more than 8 primitives"));
  std::vector<float> tris(size_t(nt) * 12), sph(size_t(ns) * 4), rects(size_t(nr) * 16);
  std::vector<int32_t> sph_mat(ns), sph_orig(ns);
  std::vector<uint32_t> storage(bvh.refs.size());  // refs position -> slot within its type's array
  std::vector<int32_t> rect_slot(nr, -1), sph_slot(ns, -1), tri_slot(nt, -1);
  uint32_t tslot = 0, sslot = 0, rslot = 0;
  for (size_t i = 0; i < bvh.refs.size(); ++i) {
    const uint32_t p = bvh.refs[i];
    if (p < nt) {
      const float* v = d->tri_v + 9 * size_t(p);
      float* o = &tris[size_t(tslot) * 12];
      int32_t m = d->tri_mat ? d->tri_mat[p] : 0;
      o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; std::memcpy(&o[3], &m, 4);
      o[4] = v[3] - v[0]; o[5] = v[4] - v[1]; o[6] = v[5] - v[2]; std::memcpy(&o[7], &p, 4);
      o[8] = v[6] - v[0]; o[9] = v[7] - v[1]; o[10] = v[8] - v[2]; o[11] = 0.0f;
      tri_slot[p] = int32_t(tslot);
      storage[i] = tslot++;
    } else if (p < nt + ns) {
      const uint32_t q = p - nt;
      std::memcpy(&sph[4 * size_t(sslot)], d->sph + 4 * size_t(q), 16);
      sph_mat[sslot] = d->sph_mat[q];
      sph_orig[sslot] = int32_t(q);
      sph_slot[q] = int32_t(sslot);
      storage[i] = sslot++;
    } else {
      const uint32_t q = p - nt - ns;
      const float* r = d->rect + 12 * size_t(q);
      float* o = &rects[16 * size_t(rslot)];
      int32_t m = d->rect_mat[q];
      o[0] = r[0]; o[1] = r[1]; o[2] = r[2]; std::memcpy(&o[3], &m, 4);
      o[4] = r[3]; o[5] = r[4]; o[6] = r[5]; o[7] = rect_area(r + 3, r + 6);
      o[8] = r[6]; o[9] = r[7]; o[10] = r[8]; std::memcpy(&o[11], &q, 4);
      o[12] = r[9]; o[13] = r[10]; o[14] = r[11]; o[15] = 0.0f;
      rect_slot[q] = int32_t(rslot);
      storage[i] = rslot++;
    }
  }
  // leaf children: refs position -> storage slot of the leaf's type
  for (size_t k = 0; k < bvh.nodes.size() / kNodeFloats; ++k) {
    float* f = &bvh.nodes[k * kNodeFloats];
    uint32_t meta;
    std::memcpy(&meta, &f[10], 4);
    for (int ch = 0; ch < 4; ++ch) {
      int32_t first;
      std::memcpy(&first, &f[12 + ch], 4);
      if (first < 0 || ((meta >> (8 * ch)) & kMetaLeaf) == 0) continue;
      int32_t slot = int32_t(storage[size_t(first)]);
      std::memcpy(&f[12 + ch], &slot, 4);
    }
  }
  // emissive primitives in the light-selection order: rectangles, spheres,
  // triangles, each in input order; codes are type << 30 | storage slot
  std::vector<uint32_t> emitters;
  auto emits = [&](int32_t m) { return d->mat[12 * size_t(m)] == float(NE_MAT_EMITTER); };
  for (uint32_t q = 0; q < nr; ++q)
    if (emits(d->rect_mat[q])) emitters.push_back(uint32_t(PRIM_RECT) << 30 | uint32_t(rect_slot[q]));
  for (uint32_t q = 0; q < ns; ++q)
    if (emits(d->sph_mat[q])) emitters.push_back(uint32_t(PRIM_SPHERE) << 30 | uint32_t(sph_slot[q]));
  for (uint32_t q = 0; q < nt; ++q)
    if (emits(d->tri_mat ? d->tri_mat[q] : 0)) emitters.push_back(uint32_t(PRIM_TRI) << 30 | uint32_t(tri_slot[q]));
  for (uint32_t e : emitters)
    if ((e & 0x3FFFFFFFu) == 0x3FFFFFFFu) return bail(fail(c, NE_ERR_INVALID_ARG, "emitter missing from the BVH"));
  std::vector<float> points(size_t(d->n_point) * 4);
  for (uint32_t q = 0; q < d->n_point; ++q) {
    for (int a = 0; a < 3; ++a) points[4 * size_t(q) + a] = d->point[3 * size_t(q) + a];
    std::memcpy(&points[4 * size_t(q) + 3], &d->point_mat[q], 4);
    emitters.push_back(uint32_t(kLightPoint) << 30 | q);
  }
  const float4* f4 = nullptr;
  if ((st = dev_upload(c, s, reinterpret_cast<const float4*>(bvh.nodes.data()), bvh.nodes.size() / 4, &f4)) != NE_OK)
    return bail(st);
  ds.nodes = f4;
  ds.n_nodes = uint32_t(bvh.nodes.size() / kNodeFloats);
  if ((st = dev_upload(c, s, reinterpret_cast<const float4*>(tris.data()), tris.size() / 4, &ds.tris)) != NE_OK)
    return bail(st);
  if ((st = dev_upload(c, s, reinterpret_cast<const float4*>(sph.data()), sph.size() / 4, &ds.sph)) != NE_OK)
    return bail(st);
  if ((st = dev_upload(c, s, sph_mat.data(), ns, &ds.sph_mat)) != NE_OK) return bail(st);
  if ((st = dev_upload(c, s, sph_orig.data(), ns, &ds.sph_orig)) != NE_OK) return bail(st);
  if ((st = dev_upload(c, s, reinterpret_cast<const float4*>(rects.data()), rects.size() / 4, &ds.rects)) != NE_OK)
    return bail(st);
  if ((st = dev_upload(c, s, d->mat, size_t(d->n_mat) * 12, &ds.mats)) != NE_OK) return bail(st);
  if ((st = dev_upload(c, s, emitters.data(), emitters.size(), &ds.emitters)) != NE_OK) return bail(st);
  if ((st = dev_upload(c, s, reinterpret_cast<const float4*>(points.data()), points.size() / 4, &ds.points)) != NE_OK)
    return bail(st);
  ds.n_emitters = int32_t(emitters.size());
  // ---- volume
  ds.has_volume = d->has_volume ? 1 : 0;
  float gmax = 0.0f;
  if (d->has_volume) {
    // A homogeneous medium is a 1^3 grid of density 1 read as constant, with
    // one majorant cell: tracking samples the exponential free flight exactly.
    static const float kOne = 1.0f;
    const bool homog = d->vol_homogeneous != 0;
    ds.homog = homog ? 1 : 0;
    for (int a = 0; a < 3; ++a) ds.res[a] = homog ? 1 : d->vol_res[a];
    for (int i = 0; i < 12; ++i) ds.inv[i] = d->vol_inv[i];
    float st3[3];
    for (int a = 0; a < 3; ++a) {
      st3[a] = d->vol_sigma_a[a] + d->vol_sigma_s[a];
      ds.sig_s[a] = d->vol_sigma_s[a];
      ds.sig_t[a] = d->vol_sigma_s[a] + d->vol_sigma_a[a];
      ds.albedo[a] = ds.sig_s[a] / ds.sig_t[a];
    }
    // avg(extinction * densityMultiplier), src/materials/GridMedia.cpp:82
    ds.sigma_bar = (st3[0] * d->vol_mult + st3[1] * d->vol_mult + st3[2] * d->vol_mult) / 3.0f;
    ds.g = d->vol_g;
    ds.mres = homog ? 1 : d->maj_res;
    {
      // row-major upload into a scratch buffer, bricked copy kept (zero padded,
      // plus one brick of slack after the last)
      const size_t nvox = size_t(ds.res[0]) * ds.res[1] * ds.res[2];
      const uint32_t px = grid_pad_x(ds.res[0]), py = grid_pad_y(ds.res[1]), pz = grid_pad_z(ds.res[2]);
      const size_t nb = (size_t(px) * py * pz + kBrickFloats) * sizeof(float);
      DevBuf lin;
      NE_HIP_BAIL(hipMalloc(&lin.p, nvox * sizeof(float)));
      NE_HIP_BAIL(hipMemcpyAsync(lin.p, homog ? &kOne : d->grid, nvox * sizeof(float), hipMemcpyHostToDevice,
                                 c->stream));
      void* pg = nullptr;
      NE_HIP_BAIL(hipMalloc(&pg, nb));
      s->allocs.push_back(pg);
      s->bytes += nb;
      NE_HIP_BAIL(hipMemsetAsync(pg, 0, nb, c->stream));
      NE_HIP_BAIL(launch_brick_grid(static_cast<const float*>(lin.p), static_cast<float*>(pg), ds.res, c->stream));
      NE_HIP_BAIL(hipStreamSynchronize(c->stream));  // before the scratch buffer is freed
      ds.grid = static_cast<const float*>(pg);
      ds.grid_step_y = (px / 4u) * kBrickFloats;
      ds.grid_step_z = (px / 4u) * (py / 4u) * kBrickFloats;
      ds.grid_small = nb < (size_t(1) << 32) && ds.grid_step_z < (1u << 24) ? 1 : 0;
    }
    size_t ncell = size_t(ds.mres) * ds.mres * ds.mres;
    void* pm = nullptr;
    NE_HIP_BAIL(hipMalloc(&pm, ncell * sizeof(float)));
    s->allocs.push_back(pm);
    s->bytes += ncell * sizeof(float);
    NE_HIP_BAIL(launch_majorants(ds, static_cast<float*>(pm), c->stream));
    ds.maj = static_cast<const float*>(pm);
    std::vector<float> mh(ncell);
    NE_HIP_BAIL(hipMemcpyAsync(mh.data(), pm, ncell * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    NE_HIP_BAIL(hipStreamSynchronize(c->stream));
    for (float v : mh) gmax = std::max(gmax, v);
    gmax = ds.sigma_bar > 0 ? gmax / ds.sigma_bar : 0.0f;
    // bounds of the cells with a positive majorant (the DDA clips rays to them)
    {
      const int M = ds.mres;
      for (int a = 0; a < 3; ++a) { ds.occ_lo[a] = M; ds.occ_hi[a] = -1; }
      for (int z = 0; z < M; ++z)
        for (int y = 0; y < M; ++y)
          for (int x = 0; x < M; ++x)
            if (mh[(size_t(z) * M + y) * M + x] > 0.0f) {
              const int cxyz[3] = {x, y, z};
              for (int a = 0; a < 3; ++a) {
                ds.occ_lo[a] = std::min(ds.occ_lo[a], cxyz[a]);
                ds.occ_hi[a] = std::max(ds.occ_hi[a], cxyz[a]);
              }
            }
    }
    if (!d->maj_no_skip && std::any_of(mh.begin(), mh.end(), [](float v) { return !(v > 0.0f); })) {
      encode_empty_distance(mh, ds.mres);
      NE_HIP_BAIL(hipMemcpyAsync(pm, mh.data(), ncell * sizeof(float), hipMemcpyHostToDevice, c->stream));
      NE_HIP_BAIL(hipStreamSynchronize(c->stream));
    }
    // the tracking loops' cell reads: the value and its reciprocal in one 8-byte load
    // (host IEEE division: the quotient the device's division returns)
    {
      std::vector<float2> mp(ncell);
      for (size_t i = 0; i < ncell; ++i) {
        const float v = mh[i];
        mp[i] = make_float2(v, v > 0.0f && std::isfinite(v) ? 1.0f / v : 0.0f);
      }
      void* pp = nullptr;
      NE_HIP_BAIL(hipMalloc(&pp, ncell * sizeof(float2)));
      s->allocs.push_back(pp);
      s->bytes += ncell * sizeof(float2);
      NE_HIP_BAIL(hipMemcpyAsync(pp, mp.data(), ncell * sizeof(float2), hipMemcpyHostToDevice, c->stream));
      NE_HIP_BAIL(hipStreamSynchronize(c->stream));
      ds.maj_pair = static_cast<const float2*>(pp);
    }
  }
  // ---- lights
  ds.has_sun = d->has_sun ? 1 : 0;
  for (int a = 0; a < 3; ++a) { ds.sun_dir[a] = d->sun_dir[a]; ds.sun_E[a] = d->sun_E[a]; }
  ds.env_type = d->env_type;
  for (int a = 0; a < 3; ++a) { ds.env[a] = d->env_c0[a]; ds.env[3 + a] = d->env_c1[a]; ds.env[6 + a] = d->env_c2[a]; }
  EnvTables et;  // outlives the asynchronous uploads below (synchronised before returning)
  if (d->env_type == NE_ENV_MAP) {
    build_env_tables(d->env_map, d->env_w, d->env_h, et);
    ds.env_w = d->env_w;
    ds.env_h = d->env_h;
    ds.env_inv_int = et.inv_int;
    if ((st = dev_upload(c, s, d->env_map, size_t(d->env_w) * d->env_h * 3, &ds.env_tex)) != NE_OK) return bail(st);
    if ((st = dev_upload(c, s, et.func.data(), et.func.size(), &ds.env_func)) != NE_OK) return bail(st);
    if ((st = dev_upload(c, s, et.cdf.data(), et.cdf.size(), &ds.env_cdf)) != NE_OK) return bail(st);
    if ((st = dev_upload(c, s, et.marg.data(), et.marg.size(), &ds.env_marg)) != NE_OK) return bail(st);
  }
  ds.n_lights = ds.has_sun + (ds.env_type != NE_ENV_NONE ? 1 : 0) + ds.n_emitters;
  NE_HIP_BAIL(hipStreamSynchronize(c->stream));
  s->info.bvh_nodes = ds.n_nodes;
  s->info.bvh_depth = bvh.max_depth;
  s->info.bvh_leaf_prims_max = bvh.max_leaf;
  s->info.maj_res = uint32_t(d->has_volume ? ds.mres : 0);
  s->info.device_bytes = s->bytes;
  s->info.grid_max_density = gmax;
  c->refs.fetch_add(1);
  *out = s;
  return NE_OK;
}
#undef NE_HIP_BAIL

void ne_scene_release(ne_scene* s) {
  if (!s) return;
  ne_ctx* c = s->ctx;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (void* p : s->allocs) (void)hipFree(p);
    delete s;
  }
  ctx_unref(c);
}

ne_status ne_scene_get_info(const ne_scene* s, ne_scene_info* out) {
  if (!s || !out) return fail(nullptr, NE_ERR_INVALID_ARG, "scene/out is NULL");
  *out = s->info;
  return NE_OK;
}

ne_status ne_film_create(ne_ctx* c, int32_t w, int32_t h, ne_film** out) {
  if (!c || !out) return fail(c, NE_ERR_INVALID_ARG, "ctx/out is NULL");
  *out = nullptr;
  if (w <= 0 || h <= 0 || w > 32768 || h > 32768) return fail(c, NE_ERR_INVALID_ARG, "film size out of range");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  ne_film* f = new ne_film();
  f->ctx = c; f->W = w; f->H = h;
  size_t bytes = size_t(w) * h * 3 * sizeof(double);
  hipError_t e = hipMalloc(&f->d_sum, bytes);
  if (e != hipSuccess) { delete f; return fail(c, NE_ERR_OUT_OF_MEMORY, "film allocation failed"); }
  e = hipMemsetAsync(f->d_sum, 0, bytes, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) { (void)hipFree(f->d_sum); delete f; return fail(c, NE_ERR_HIP, "film clear failed"); }
  c->refs.fetch_add(1);
  *out = f;
  return NE_OK;
}

void ne_film_release(ne_film* f) {
  if (!f) return;
  ne_ctx* c = f->ctx;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(f->d_sum);
    if (f->d_fixed) (void)hipFree(f->d_fixed);
    delete f;
  }
  ctx_unref(c);
}

ne_status ne_film_clear(ne_film* f) {
  if (!f) return fail(nullptr, NE_ERR_INVALID_ARG, "film is NULL");
  ne_ctx* c = f->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  NE_HIP(c, hipMemsetAsync(f->d_sum, 0, size_t(f->W) * f->H * 3 * sizeof(double), c->stream));
  return NE_OK;
}

ne_status ne_film_read(ne_film* f, double* sum) {
  if (!f || !sum) return fail(f ? f->ctx : nullptr, NE_ERR_INVALID_ARG, "film/sum is NULL");
  ne_ctx* c = f->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  NE_HIP(c, hipMemcpyAsync(sum, f->d_sum, size_t(f->W) * f->H * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

ne_status ne_film_tonemap(ne_film* f, int32_t spp, float* pixels) {
  if (!f || !pixels || spp <= 0) return fail(f ? f->ctx : nullptr, NE_ERR_INVALID_ARG, "bad tonemap arguments");
  ne_ctx* c = f->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  const int npix = f->W * f->H;
  DevBuf b;
  ne_status st = tmp_out<float>(c, b, size_t(npix) * 3);
  if (st != NE_OK) return st;
  NE_HIP(c, launch_tonemap(f->d_sum, npix, 1.0f / float(spp), static_cast<float*>(b.p), c->stream));
  st = tmp_back(c, pixels, b, size_t(npix) * 3);
  if (st != NE_OK) return st;
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

constexpr size_t kMaxSampleBuffer = size_t(16) << 30;

// Default wavefront pool (slots) for a render of `total` paths, measured on
// MI355X: 2^23 for surface-only scenes, where fewer, larger BVH launches
// amortise each launch's tail; with a volume 2^22, where a larger pool adds
// cache pressure, except for long renders (2^28 paths and more, e.g. the
// 1920x1080 cloud at 256 spp), whose iterations' fixed costs (launches, the
// plan scan) halve with 2^23 slots: 236.4 -> 231.5 ms per cloud frame, while
// a 512^2 x 64-spp volume frame took 12.0 ms at 2^22 and 13.4 ms at 2^23.
static uint64_t default_pool(const DevScene& ds, uint64_t total) {
  if (ds.n_nodes > 0 && !ds.has_volume) return uint64_t(1) << 23;
  return total >= (uint64_t(1) << 28) ? (uint64_t(1) << 23) : (uint64_t(1) << 22);
}

static ne_status render_wavefront(ne_ctx* c, const ne_scene* s, const DevCamera& dc, const ne_render_params* p,
                                  ne_film* f, bool count, bool timed, StageTimer* tm, WfGuideHook* gh = nullptr) {
  const uint64_t total = uint64_t(p->x1 - p->x0) * uint64_t(p->y1 - p->y0) * uint64_t(p->spp);
  // pool slots: whole chunks, no more than the job needs
  uint64_t P = p->pool_size > 0 ? uint64_t(p->pool_size) : default_pool(s->dev, total);
  P = std::min<uint64_t>(P, total);
  P = (P + kChunk - 1) / kChunk * kChunk;
  if (P > c->wf_P) {
    if (c->wf_buf) {
      NE_HIP(c, hipStreamSynchronize(c->stream));
      void* old = c->wf_buf;  // cleared first: a failing free never leaves it to be freed twice
      c->wf_buf = nullptr;
      c->wf_P = 0;
      NE_HIP(c, hipFree(old));
    }
    void* nb = nullptr;
    NE_HIP(c, hipMalloc(&nb, wf_bytes(uint32_t(P))));
    c->wf_buf = nb;
    c->wf_P = uint32_t(P);
  }
  if (!f->d_fixed) {
    const size_t bytes = size_t(f->W) * f->H * 3 * sizeof(long long);
    NE_HIP(c, hipMalloc(&f->d_fixed, bytes));
    NE_HIP(c, hipMemsetAsync(f->d_fixed, 0, bytes, c->stream));
  }
  WfState w{};
  wf_carve(c->wf_buf, uint32_t(P), w);
  WfArgs a{};
  a.W = p->width; a.H = p->height; a.x0 = p->x0; a.y0 = p->y0; a.x1 = p->x1; a.y1 = p->y1;
  a.max_bounces = p->max_bounces; a.sample_offset = p->sample_offset; a.job = p->job; a.nchunks = uint32_t(P / kChunk);
  a.seed = p->seed;
  a.sort = ((p->flags & NE_FLAG_SORT) ? 1 : 0) | ((p->flags & NE_FLAG_SPATIAL) ? 2 : 0);
  a.wide_offsets = (p->flags & NE_FLAG_WIDE_OFFSETS) ? 1 : 0;
  a.plan_stall = (p->flags & NE_FLAG_TEST_PLAN_STALL) ? 1u : 0u;
  a.refill_min = kRefillMinIdle;
  if (c->refill_blocks == 0) {
    int cus = 0, per_cu = 0;
    NE_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    NE_HIP(c, wf_refill_blocks_per_cu(&per_cu));
    c->refill_blocks = uint32_t(std::max(1, cus * std::max(1, per_cu)));
  }
  // chunks per refill block: one round of resident blocks covers the pool
  a.refill = std::min<uint32_t>(kMaxRefill, std::max<uint32_t>(1, (a.nchunks + c->refill_blocks - 1) / c->refill_blocks));
  a.total_paths = total;
  a.counters = c->d_counters;
  a.film_fixed = f->d_fixed;
  a.film = f->d_sum;
  // Per-sample radiance slots replace the film atomics when the render's
  // samples fit in kMaxSampleBuffer (12 B each; 6.4 GB for 1920x1080 at 256
  // spp); larger renders use the fixed-point film.
  a.sbuf = nullptr;
  const size_t sbytes = size_t(total) * 3 * sizeof(float);
  if (sbytes <= kMaxSampleBuffer && !(p->flags & NE_FLAG_FIXED_FILM)) {
    if (sbytes > c->sbuf_bytes) {
      NE_HIP(c, hipStreamSynchronize(c->stream));
      void* old = c->sbuf;
      c->sbuf = nullptr;
      c->sbuf_bytes = 0;
      if (old) NE_HIP(c, hipFree(old));
      void* nb = nullptr;
      if (hipMalloc(&nb, sbytes) == hipSuccess) {
        c->sbuf = nb;
        c->sbuf_bytes = sbytes;
      } else {
        (void)hipGetLastError();  // no room: the fixed-point film takes this render
      }
    }
    a.sbuf = sbytes <= c->sbuf_bytes ? static_cast<float*>(c->sbuf) : nullptr;
  }
  a.gstate[0] = a.gstate[1] = nullptr;
  if (gh) {
    if (P > c->wfg_P) {
      if (c->wfg_buf) {
        NE_HIP(c, hipStreamSynchronize(c->stream));
        void* old = c->wfg_buf;
        c->wfg_buf = nullptr;
        c->wfg_P = 0;
        NE_HIP(c, hipFree(old));
      }
      void* nb = nullptr;
      NE_HIP(c, hipMalloc(&nb, wf_guide_bytes(uint32_t(P))));
      c->wfg_buf = nb;
      c->wfg_P = uint32_t(P);
    }
    wf_guide_carve(c->wfg_buf, uint32_t(P), gh->g);
    a.gstate[0] = gh->g.rec[0].state;
    a.gstate[1] = gh->g.rec[1].state;
  }
  if (timed) NE_HIP(c, hipEventRecord(c->ev0, c->stream));
  NE_HIP(c, launch_wavefront(s->dev, dc, w, a, count, c->stream, c->h_status, tm, gh));
  if (timed) {
    NE_HIP(c, hipEventRecord(c->ev1, c->stream));
    c->timed = true;
  }
  return NE_OK;
}

// Every argument check of a render call, made before anything touches the film.
static ne_status check_render(ne_ctx* c, const ne_scene* s, const ne_camera_desc* cam, const ne_render_params* p,
                              const ne_film* f) {
  ne_status st = check_params(c, p);
  if (st != NE_OK) return st;
  if (!s || !cam || !f) return fail(c, NE_ERR_INVALID_ARG, "scene/camera/film is NULL");
  if (s->ctx != c || f->ctx != c) return fail(c, NE_ERR_INVALID_ARG, "scene/film belong to another context");
  if (f->W != p->width || f->H != p->height) return fail(c, NE_ERR_INVALID_ARG, "film size differs from params");
  return NE_OK;
}

// Megakernel launch arguments: one thread per (pixel, chunk of S samples); S
// from the params or sized for about 2^22 threads.
static RenderArgs megakernel_args(ne_ctx* c, const ne_render_params* p) {
  RenderArgs a{};
  a.W = p->width; a.H = p->height; a.x0 = p->x0; a.y0 = p->y0; a.x1 = p->x1; a.y1 = p->y1;
  a.spp = p->spp; a.max_bounces = p->max_bounces; a.sample_offset = p->sample_offset; a.job = p->job;
  a.seed = p->seed;
  const int64_t npix = int64_t(p->x1 - p->x0) * (p->y1 - p->y0);
  int S = p->samples_per_thread;
  if (S <= 0) {
    const int64_t target_threads = int64_t(1) << 22;
    int64_t want = (npix * p->spp + target_threads - 1) / target_threads;
    S = int(std::min<int64_t>(std::max<int64_t>(want, 1), p->spp));
  }
  S = std::min(S, p->spp);
  a.S = S;
  a.nchunks = (p->spp + S - 1) / S;
  a.ntx = (p->x1 - p->x0 + 7) / 8;
  a.nty = (p->y1 - p->y0 + 7) / 8;
  a.counters = c->d_counters;
  return a;
}

static ne_status render_locked(ne_ctx* c, const ne_scene* s, const ne_camera_desc* cam, const ne_render_params* p,
                               ne_film* f, const GuideUse* gu = nullptr) {
  ne_status st = check_render(c, s, cam, p, f);
  if (st != NE_OK) return st;
  NE_HIP(c, hipSetDevice(c->device));
  const bool count = (p->flags & NE_FLAG_COUNT) != 0;
  const bool timed = (p->flags & NE_FLAG_TIME) != 0;
  if (count) {
    NE_HIP(c, hipMemsetAsync(c->d_counters, 0, sizeof(DevCounters), c->stream));
    c->counters_valid = true;
  }
  DevCamera dc = make_camera(cam);
  StageTimer timer{&c->ev_pool, &c->ev_stages, c->stream};
  StageTimer* tm = nullptr;
  c->stage_timed = false;
  if (p->flags & NE_FLAG_STAGE_TIME) {
    c->ev_stages.clear();
    tm = &timer;
  }
  const bool wavefront = p->mode == NE_MODE_WAVEFRONT || (p->mode == NE_MODE_AUTO && p->precision == NE_FP32);
  if (wavefront) {
    st = render_wavefront(c, s, dc, p, f, count, timed, tm);
    c->stage_timed = st == NE_OK && tm;
    return st;
  }
  RenderArgs a = megakernel_args(c, p);
  if (gu) a.guide = *gu;
  const int64_t npix = int64_t(p->x1 - p->x0) * (p->y1 - p->y0);
  if (a.nchunks == 1) {
    a.out = f->d_sum;
  } else {
    size_t need = size_t(npix) * size_t(a.nchunks) * 3 * sizeof(double);
    if (need > c->partial_bytes) {
      if (c->d_partial) {
        NE_HIP(c, hipStreamSynchronize(c->stream));
        void* old = c->d_partial;
        c->d_partial = nullptr;
        c->partial_bytes = 0;
        NE_HIP(c, hipFree(old));
      }
      void* nb = nullptr;
      NE_HIP(c, hipMalloc(&nb, need));
      c->d_partial = static_cast<decltype(c->d_partial)>(nb);
      c->partial_bytes = need;
    }
    a.out = c->d_partial;
  }
  int block = p->block_size ? p->block_size : 128;
  if (timed) NE_HIP(c, hipEventRecord(c->ev0, c->stream));
  if (tm) NE_HIP(c, tm->mark(kStageMegakernel));
  NE_HIP(c, launch_render(s->dev, dc, a, p->precision, count, block, p->occupancy, c->stream));
  if (timed) {
    NE_HIP(c, hipEventRecord(c->ev1, c->stream));
    c->timed = true;
  }
  if (a.nchunks > 1) {
    if (tm) NE_HIP(c, tm->mark(kStageOther));
    NE_HIP(c, launch_reduce_partials(a, f->d_sum, c->stream));
  }
  if (tm) {
    NE_HIP(c, tm->mark(kStageEnd));
    c->stage_timed = true;
  }
  return NE_OK;
}

ne_status ne_render(ne_ctx* c, const ne_scene* s, const ne_camera_desc* cam, const ne_render_params* p, ne_film* f) {
  if (!c) return fail(nullptr, NE_ERR_INVALID_ARG, "ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  return render_locked(c, s, cam, p, f);
}

ne_status ne_render_noalloc(ne_ctx* c, const ne_scene* s, const ne_camera_desc* cam, const ne_render_params* p,
                           ne_film* f) {
  if (!c) return fail(nullptr, NE_ERR_INVALID_ARG, "ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  ne_status st = check_render(c, s, cam, p, f);
  if (st != NE_OK) return st;
  if (p->mode == NE_MODE_WAVEFRONT) return fail(c, NE_ERR_INVALID_ARG, "ne_render_noalloc runs the megakernel only");
  if (p->flags & (NE_FLAG_COUNT | NE_FLAG_STAGE_TIME))
    return fail(c, NE_ERR_INVALID_ARG, "ne_render_noalloc does not count events or time stages");
  RenderArgs a = megakernel_args(c, p);
  const size_t npix = size_t(p->x1 - p->x0) * size_t(p->y1 - p->y0);
  if (a.nchunks == 1) {
    a.out = f->d_sum;  // each thread adds its pixel's sums into the film: no scratch
  } else {
    const size_t need = npix * size_t(a.nchunks) * 3 * sizeof(double);
    if (need > kStaticScratchBytes)
      return fail(c, NE_ERR_INVALID_ARG,
                  "partial sums need " + std::to_string(need) + " bytes, the static scratch holds " +
                      std::to_string(kStaticScratchBytes) + " (render fewer pixels or raise samples_per_thread)");
    // every (chunk, pixel) slot the reduction reads is written by the render
    // first, so the previous call's contents need no clearing
    a.out = c->scratch;
  }
  NE_HIP(c, hipSetDevice(c->device));
  const bool timed = (p->flags & NE_FLAG_TIME) != 0;
  if (timed) NE_HIP(c, hipEventRecord(c->ev0, c->stream));
  NE_HIP(c, launch_render(s->dev, make_camera(cam), a, p->precision, false, p->block_size ? p->block_size : 128,
                          p->occupancy, c->stream));
  if (timed) {
    NE_HIP(c, hipEventRecord(c->ev1, c->stream));
    c->timed = true;
  }
  if (a.nchunks > 1) NE_HIP(c, launch_reduce_partials(a, f->d_sum, c->stream));
  return NE_OK;
}

ne_status ne_render_tile(ne_ctx* c, const ne_scene* s, const ne_camera_desc* cam, const ne_render_params* p,
                         int32_t index, ne_film* f, float* pixels) {
  if (!c || !p || !pixels || !f) return fail(c, NE_ERR_INVALID_ARG, "NULL argument");
  // OfflineEngine tiling: tileSize = resolution / (40, 10), tile (index % 40, index / 40)
  const int tw = p->width / 40, th = p->height / 10;
  if (index < 0 || index >= 400) return fail(c, NE_ERR_INVALID_ARG, "tile index out of range [0, 400)");
  std::lock_guard<std::mutex> lk(c->mu);
  if (f->ctx != c) return fail(c, NE_ERR_INVALID_ARG, "scene/film belong to another context");
  if (f->W != p->width || f->H != p->height) return fail(c, NE_ERR_INVALID_ARG, "film size differs from params");
  // Below 40 x 10 pixels the integer tile size is 0 on an axis and renderTile's
  // loops run zero times: the tile is empty, nothing is rendered or written.
  ne_render_params q = *p;
  if (tw <= 0 || th <= 0) {
    q.x0 = 0; q.y0 = 0; q.x1 = p->width; q.y1 = p->height;  // the tile rectangle is not used
    return check_params(c, &q);
  }
  q.x0 = (index % 40) * tw; q.y0 = (index / 40) * th; q.x1 = q.x0 + tw; q.y1 = q.y0 + th;
  // a rejected call leaves the film as it was
  ne_status st = check_render(c, s, cam, &q, f);
  if (st != NE_OK) return st;
  NE_HIP(c, hipSetDevice(c->device));
  // renderTile recomputes the tile: its film sums restart from zero
  NE_HIP(c, launch_clear_rect(f->d_sum, f->W, q.x0, q.y0, q.x1, q.y1, c->stream));
  st = render_locked(c, s, cam, &q, f);
  if (st != NE_OK) return st;
  // post-process the tile only, then one strided copy into the caller's
  // full-frame pixel array
  DevBuf b;
  if ((st = tmp_out<float>(c, b, size_t(tw) * th * 3)) != NE_OK) return st;
  NE_HIP(c, launch_tonemap_rect(f->d_sum, f->W, q.x0, q.y0, tw, th, 1.0f / float(p->spp), static_cast<float*>(b.p),
                                c->stream));
  const size_t row = size_t(tw) * 3 * sizeof(float);
  NE_HIP(c, hipMemcpy2DAsync(pixels + (size_t(q.y0) * f->W + q.x0) * 3, size_t(f->W) * 3 * sizeof(float), b.p, row,
                             row, size_t(th), hipMemcpyDeviceToHost, c->stream));
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

ne_status ne_get_kernel_time(ne_ctx* c, float* ms) {
  if (!c || !ms) return fail(c, NE_ERR_INVALID_ARG, "ctx/ms is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->timed) return fail(c, NE_ERR_INVALID_ARG, "no render with NE_FLAG_TIME has been issued");
  NE_HIP(c, hipSetDevice(c->device));
  NE_HIP(c, hipEventSynchronize(c->ev1));
  NE_HIP(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
  return NE_OK;
}

ne_status ne_get_stage_times(ne_ctx* c, float* ms, uint32_t* launches) {
  if (!c || !ms || !launches) return fail(c, NE_ERR_INVALID_ARG, "ctx/ms/launches is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->stage_timed || c->ev_stages.empty())
    return fail(c, NE_ERR_INVALID_ARG, "no render with NE_FLAG_STAGE_TIME has been issued");
  NE_HIP(c, hipSetDevice(c->device));
  for (int k = 0; k < NE_STAGE_COUNT; ++k) {
    ms[k] = 0.0f;
    launches[k] = 0;
  }
  const size_t n = c->ev_stages.size();
  NE_HIP(c, hipEventSynchronize(c->ev_pool[n - 1]));
  for (size_t k = 0; k + 1 < n; ++k) {
    const int stage = c->ev_stages[k];
    if (stage < 0 || stage >= NE_STAGE_COUNT) continue;
    float t = 0.0f;
    NE_HIP(c, hipEventElapsedTime(&t, c->ev_pool[k], c->ev_pool[k + 1]));
    ms[stage] += t;
    launches[stage] += 1;
  }
  return NE_OK;
}

ne_status ne_get_counters(ne_ctx* c, ne_counters* out) {
  if (!c || !out) return fail(c, NE_ERR_INVALID_ARG, "ctx/out is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->counters_valid) return fail(c, NE_ERR_INVALID_ARG, "no render with NE_FLAG_COUNT has been issued");
  NE_HIP(c, hipSetDevice(c->device));
  DevCounters h{};
  NE_HIP(c, hipMemcpyAsync(&h, c->d_counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  NE_HIP(c, hipStreamSynchronize(c->stream));
  out->paths = h.paths; out->ext_rays = h.ext_rays; out->shadow_rays = h.shadow_rays;
  // device fields hold the extension-ray share; totals add the shadow share
  out->density_lookups = h.density_lookups + h.sh_lookups; out->majorant_cells = h.majorant_cells + h.sh_cells;
  out->bvh_nodes = h.bvh_nodes + h.sh_nodes; out->prim_tests = h.prim_tests + h.sh_prims;
  out->shadow_density_lookups = h.sh_lookups; out->shadow_majorant_cells = h.sh_cells;
  out->shadow_bvh_nodes = h.sh_nodes; out->shadow_prim_tests = h.sh_prims;
  out->guide_queries = h.guide_queries;
  return NE_OK;
}

ne_status ne_li(ne_ctx* c, const ne_scene* s, int32_t n, const float* rays, int32_t max_bounces, uint64_t seed,
                uint32_t sample_index, float* radiance) {
  if (!c || !s || n < 0 || (n > 0 && (!rays || !radiance))) return fail(c, NE_ERR_INVALID_ARG, "bad ne_li arguments");
  // the shadow-ray sub-streams of a path are keyed by vertex below 2^11: a
  // longer path would run them into the main stream's counter range
  if (max_bounces < 0 || max_bounces > kMaxVertices)
    return fail(c, NE_ERR_INVALID_ARG, "max_bounces out of range [0, 2047]");
  if (s->ctx != c) return fail(c, NE_ERR_INVALID_ARG, "scene belongs to another context");
  if (n == 0) return NE_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  DevBuf br, bo;
  ne_status st;
  if ((st = tmp_in(c, br, rays, size_t(n) * 6)) != NE_OK) return st;
  if ((st = tmp_out<float>(c, bo, size_t(n) * 3)) != NE_OK) return st;
  NE_HIP(c, launch_li(s->dev, n, static_cast<const float*>(br.p), max_bounces, seed, sample_index,
                      static_cast<float*>(bo.p), c->stream));
  if ((st = tmp_back(c, radiance, bo, size_t(n) * 3)) != NE_OK) return st;
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

// ---- parity entry points ---------------------------------------------------
ne_status ne_test_primitive(ne_ctx* c, int32_t kind, int32_t n, const float* rays, const float* prims, int32_t* hit,
                            float* t0, float* t1) {
  if (!c || kind < 0 || kind > 3 || n < 0) return fail(c, NE_ERR_INVALID_ARG, "bad ne_test_primitive arguments");
  if (n == 0) return NE_OK;
  if (!rays || !prims || !hit || !t0 || !t1) return fail(c, NE_ERR_INVALID_ARG, "NULL buffer");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  const size_t pw = kind == 0 ? 6 : (kind == 2 ? 4 : 9);
  DevBuf br, bp, bh, b0, b1;
  ne_status st;
  if ((st = tmp_in(c, br, rays, size_t(n) * 6)) || (st = tmp_in(c, bp, prims, size_t(n) * pw)) ||
      (st = tmp_out<int32_t>(c, bh, n)) || (st = tmp_out<float>(c, b0, n)) || (st = tmp_out<float>(c, b1, n)))
    return st;
  NE_HIP(c, launch_test_primitive(kind, n, (const float*)br.p, (const float*)bp.p, (int32_t*)bh.p, (float*)b0.p,
                                  (float*)b1.p, c->stream));
  if ((st = tmp_back(c, hit, bh, n)) || (st = tmp_back(c, t0, b0, n)) || (st = tmp_back(c, t1, b1, n))) return st;
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

ne_status ne_test_closest(ne_ctx* c, const ne_scene* s, int32_t n, const float* rays, float tmin, float tmax, float* t,
                          int32_t* type, int32_t* index) {
  if (!c || !s || n < 0) return fail(c, NE_ERR_INVALID_ARG, "bad ne_test_closest arguments");
  if (n == 0) return NE_OK;
  if (!rays || !t || !type || !index) return fail(c, NE_ERR_INVALID_ARG, "NULL buffer");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  DevBuf br, bt, by, bi;
  ne_status st;
  if ((st = tmp_in(c, br, rays, size_t(n) * 6)) || (st = tmp_out<float>(c, bt, n)) ||
      (st = tmp_out<int32_t>(c, by, n)) || (st = tmp_out<int32_t>(c, bi, n)))
    return st;
  NE_HIP(c, launch_test_closest(s->dev, n, (const float*)br.p, tmin, tmax, (float*)bt.p, (int32_t*)by.p,
                                (int32_t*)bi.p, c->stream));
  if ((st = tmp_back(c, t, bt, n)) || (st = tmp_back(c, type, by, n)) || (st = tmp_back(c, index, bi, n))) return st;
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

ne_status ne_test_transmittance(ne_ctx* c, const ne_scene* s, int32_t n, const float* rays, const float* tmax,
                                uint64_t seed, int32_t precision, double* tr) {
  if (!c || !s || n < 0 || (precision != 0 && precision != 1)) return fail(c, NE_ERR_INVALID_ARG, "bad arguments");
  if (n == 0) return NE_OK;
  if (!rays || !tmax || !tr) return fail(c, NE_ERR_INVALID_ARG, "NULL buffer");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  DevBuf br, bm, bo;
  ne_status st;
  if ((st = tmp_in(c, br, rays, size_t(n) * 6)) || (st = tmp_in(c, bm, tmax, n)) || (st = tmp_out<double>(c, bo, n)))
    return st;
  NE_HIP(c, launch_test_transmittance(s->dev, n, (const float*)br.p, (const float*)bm.p, seed, precision,
                                      (double*)bo.p, c->stream));
  if ((st = tmp_back(c, tr, bo, n))) return st;
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

ne_status ne_test_delta_track(ne_ctx* c, const ne_scene* s, int32_t n, const float* rays, uint64_t seed,
                              int32_t precision, int32_t* scattered, double* t) {
  if (!c || !s || n < 0 || (precision != 0 && precision != 1)) return fail(c, NE_ERR_INVALID_ARG, "bad arguments");
  if (n == 0) return NE_OK;
  if (!rays || !scattered || !t) return fail(c, NE_ERR_INVALID_ARG, "NULL buffer");
  if (!s->dev.has_volume) return fail(c, NE_ERR_INVALID_ARG, "scene has no volume");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  DevBuf br, bs, bt;
  ne_status st;
  if ((st = tmp_in(c, br, rays, size_t(n) * 6)) || (st = tmp_out<int32_t>(c, bs, n)) || (st = tmp_out<double>(c, bt, n)))
    return st;
  NE_HIP(c, launch_test_delta(s->dev, n, (const float*)br.p, seed, precision, (int32_t*)bs.p, (double*)bt.p,
                              c->stream));
  if ((st = tmp_back(c, scattered, bs, n)) || (st = tmp_back(c, t, bt, n))) return st;
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

ne_status ne_test_density(ne_ctx* c, const ne_scene* s, int32_t n, const float* gcs, float* out) {
  if (!c || !s || n < 0) return fail(c, NE_ERR_INVALID_ARG, "bad arguments");
  if (n == 0) return NE_OK;
  if (!gcs || !out) return fail(c, NE_ERR_INVALID_ARG, "NULL buffer");
  if (!s->dev.has_volume) return fail(c, NE_ERR_INVALID_ARG, "scene has no volume");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  DevBuf bg, bo;
  ne_status st;
  if ((st = tmp_in(c, bg, gcs, size_t(n) * 3)) || (st = tmp_out<float>(c, bo, n))) return st;
  NE_HIP(c, launch_test_density(s->dev, n, (const float*)bg.p, (float*)bo.p, c->stream));
  if ((st = tmp_back(c, out, bo, n))) return st;
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

ne_status ne_test_majorants(ne_ctx* c, const ne_scene* s, float* out, int64_t capacity) {
  if (!c || !s || !out) return fail(c, NE_ERR_INVALID_ARG, "bad arguments");
  if (!s->dev.has_volume) return fail(c, NE_ERR_INVALID_ARG, "scene has no volume");
  const int64_t M = s->dev.mres, n = M * M * M;
  if (capacity < n) return fail(c, NE_ERR_INVALID_ARG, "output holds fewer than maj_res^3 floats");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  NE_HIP(c, hipMemcpyAsync(out, s->dev.maj, size_t(n) * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

ne_status ne_test_bsdf(ne_ctx* c, int32_t kind, int32_t n, const float* in0, const float* in1, const float* in2,
                       const float* in3, float* out) {
  if (!c || kind < 0 || kind > 5 || n < 0) return fail(c, NE_ERR_INVALID_ARG, "bad ne_test_bsdf arguments");
  if (n == 0) return NE_OK;
  // element widths of (in0, in1, in2, in3, out) per kind
  static const int w[6][5] = {{3, 3, 1, 0, 1}, {2, 1, 0, 0, 3}, {3, 1, 0, 0, 1},
                              {3, 3, 1, 0, 1}, {5, 3, 3, 3, 7}, {3, 2, 1, 0, 4}};
  const float* ins[4] = {in0, in1, in2, in3};
  for (int k = 0; k < 4; ++k)
    if (w[kind][k] && !ins[k]) return fail(c, NE_ERR_INVALID_ARG, "NULL input buffer");
  if (!out) return fail(c, NE_ERR_INVALID_ARG, "NULL output buffer");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  DevBuf b[4], bo;
  ne_status st;
  for (int k = 0; k < 4; ++k)
    if (w[kind][k] && (st = tmp_in(c, b[k], ins[k], size_t(n) * w[kind][k])) != NE_OK) return st;
  if ((st = tmp_out<float>(c, bo, size_t(n) * w[kind][4]))) return st;
  NE_HIP(c, launch_test_bsdf(kind, n, (const float*)b[0].p, (const float*)b[1].p, (const float*)b[2].p,
                             (const float*)b[3].p, (float*)bo.p, c->stream));
  if ((st = tmp_back(c, out, bo, size_t(n) * w[kind][4]))) return st;
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

ne_status ne_test_rng(ne_ctx* c, uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t job, int32_t n, float* out) {
  if (!c || n < 0 || (n > 0 && !out)) return fail(c, NE_ERR_INVALID_ARG, "bad ne_test_rng arguments");
  if (n == 0) return NE_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  DevBuf bo;
  ne_status st;
  if ((st = tmp_out<float>(c, bo, n))) return st;
  NE_HIP(c, launch_test_rng(seed, pixel, sample, job, n, (float*)bo.p, c->stream));
  if ((st = tmp_back(c, out, bo, n))) return st;
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

}  // extern "C"

// ---- batched rendering -------------------------------------------------------
// A stream of the wavefront executor, with its own path pool and status word.
struct BatchLane {
  hipStream_t stream = nullptr;
  void* pool = nullptr;
  uint32_t pool_P = 0;
  void* sbuf = nullptr;  // per-sample radiance slots (12 B per sample of the lane's largest job)
  size_t sbuf_bytes = 0;
  uint32_t* status = nullptr;  // pinned host words read by the wavefront loop
};

struct ne_batch {
  ne_ctx* ctx = nullptr;
  std::vector<ne_job> jobs;
  std::vector<size_t> film_off;   // per job, in long long words
  size_t film_words = 0;
  long long* d_films = nullptr;
  BatchJob* d_jobs = nullptr;
  unsigned long long* d_misc = nullptr;  // [0] queue counter, [1] start time
  float* d_job_ms = nullptr;
  unsigned long long* d_item_time = nullptr;
  size_t item_cap = 0;
  ne_batch_config tuned{8, 8, 4, 0};
  hipGraphExec_t exec = nullptr;
  ne_batch_config exec_cfg{};
  int exec_div = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<float> job_ms;
  uint32_t grid_blocks = 0;
  double ticks_per_ms = 0;
  bool surfaces = true;  // some job has surfaces or emitters (the general batch kernel)
  bool guided = false;   // some job samples a guide's tabulated proposal
  std::vector<BatchLane> lanes;  // wavefront executor (created on its first run)
};

namespace {

void batch_free(ne_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->d_films) (void)hipFree(b->d_films);
  if (b->d_jobs) (void)hipFree(b->d_jobs);
  if (b->d_misc) (void)hipFree(b->d_misc);
  if (b->d_job_ms) (void)hipFree(b->d_job_ms);
  if (b->d_item_time) (void)hipFree(b->d_item_time);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  for (BatchLane& L : b->lanes) {
    if (L.stream) (void)hipStreamDestroy(L.stream);
    if (L.pool) (void)hipFree(L.pool);
    if (L.sbuf) (void)hipFree(L.sbuf);
    if (L.status) (void)hipHostFree(L.status);
  }
  b->lanes.clear();
}

bool cfg_valid(const ne_batch_config& c) {
  if (c.executor == NE_BATCH_WAVEFRONT) return true;  // no item shape
  return c.executor == NE_BATCH_PERSISTENT && c.tile_w >= 1 && c.tile_h >= 1 && c.tile_w * c.tile_h <= 4096 &&
         c.samples_per_item >= 1 && c.samples_per_item <= 65536;
}

// Builds the job table for an item shape and sample divisor and (re)captures
// the launch sequence into a graph when either changed.
ne_status batch_prepare(ne_batch* b, const ne_batch_config& cfg, int div) {
  ne_ctx* c = b->ctx;
  std::vector<BatchJob> table(b->jobs.size());
  unsigned long long items = 0;
  for (size_t j = 0; j < b->jobs.size(); ++j) {
    const ne_job& J = b->jobs[j];
    BatchJob& T = table[j];
    T.scene = J.scene->dev;
    T.cam = make_camera(&J.camera);
    T.W = J.width; T.H = J.height;
    T.spp = (J.spp + div - 1) / div;
    T.max_bounces = J.max_bounces;
    T.seed = J.seed;
    T.job = J.job;
    T.tiles_x = uint32_t((J.width + cfg.tile_w - 1) / cfg.tile_w);
    T.tiles_y = uint32_t((J.height + cfg.tile_h - 1) / cfg.tile_h);
    T.chunks = uint32_t((T.spp + cfg.samples_per_item - 1) / cfg.samples_per_item);
    T.first_item = items;
    T.film = b->d_films + b->film_off[j];
    if (J.guide) T.guide = guide_use(J.guide);
    items += (unsigned long long)T.tiles_x * T.tiles_y * T.chunks;
  }
  if (items > b->item_cap) {
    NE_HIP(c, hipStreamSynchronize(c->stream));
    // handles are cleared before the calls that release them, so a failing
    // call never leaves one to be released twice
    void* old_times = b->d_item_time;
    b->d_item_time = nullptr;
    b->item_cap = 0;
    hipGraphExec_t old_exec = b->exec;
    b->exec = nullptr;
    if (old_times) NE_HIP(c, hipFree(old_times));
    if (old_exec) NE_HIP(c, hipGraphExecDestroy(old_exec));
    void* p = nullptr;
    NE_HIP(c, hipMalloc(&p, items * sizeof(unsigned long long)));
    b->d_item_time = static_cast<unsigned long long*>(p);
    b->item_cap = items;
  }
  NE_HIP(c, hipMemcpyAsync(b->d_jobs, table.data(), table.size() * sizeof(BatchJob), hipMemcpyHostToDevice,
                           c->stream));
  const bool same = b->exec && b->exec_div == div && b->exec_cfg.tile_w == cfg.tile_w &&
                    b->exec_cfg.tile_h == cfg.tile_h && b->exec_cfg.samples_per_item == cfg.samples_per_item;
  if (same) return NE_OK;
  if (b->exec) {
    NE_HIP(c, hipStreamSynchronize(c->stream));
    hipGraphExec_t old_exec = b->exec;
    b->exec = nullptr;
    NE_HIP(c, hipGraphExecDestroy(old_exec));
  }
  BatchArgs a{};
  a.jobs = b->d_jobs;
  a.njobs = int32_t(b->jobs.size());
  a.tile_w = uint32_t(cfg.tile_w); a.tile_h = uint32_t(cfg.tile_h); a.spi = uint32_t(cfg.samples_per_item);
  a.total_items = items;
  a.next_item = b->d_misc;
  a.start_time = b->d_misc + 1;
  a.item_time = b->d_item_time;
  a.job_ms = b->d_job_ms;
  a.films = b->d_films;
  // the job table copy above is ordinary stream work; capture starts after it
  NE_HIP(c, hipStreamSynchronize(c->stream));
  hipGraph_t graph = nullptr;
  NE_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = launch_batch(a, b->surfaces, b->guided, b->film_words, b->grid_blocks, b->ticks_per_ms, c->stream);
  hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(c, NE_ERR_HIP, std::string("batch graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
  }
  e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    b->exec = nullptr;
    return fail(c, NE_ERR_HIP, std::string("batch graph instantiation failed: ") + hipGetErrorString(e));
  }
  b->exec_cfg = cfg;
  b->exec_div = div;
  return NE_OK;
}

ne_status batch_launch(ne_batch* b, const ne_batch_config& cfg, int div, float* ms) {
  ne_ctx* c = b->ctx;
  ne_status st = batch_prepare(b, cfg, div);
  if (st != NE_OK) return st;
  // guided jobs read their guide's current tabulation
  for (const ne_job& J : b->jobs)
    if (J.guide && (st = guide_refresh_cache(c, const_cast<ne_guide*>(J.guide))) != NE_OK) return st;
  NE_HIP(c, hipEventRecord(b->ev0, c->stream));
  NE_HIP(c, hipGraphLaunch(b->exec, c->stream));
  NE_HIP(c, hipEventRecord(b->ev1, c->stream));
  NE_HIP(c, hipEventSynchronize(b->ev1));
  float t = 0;
  NE_HIP(c, hipEventElapsedTime(&t, b->ev0, b->ev1));
  if (ms) *ms = t;
  b->job_ms.assign(b->jobs.size(), 0.0f);
  NE_HIP(c, hipMemcpy(b->job_ms.data(), b->d_job_ms, b->jobs.size() * sizeof(float), hipMemcpyDeviceToHost));
  return NE_OK;
}

// Path-pool slots and sample count of job j at ceil(spp / div) samples, as
// ne_render sizes them.
static uint64_t lane_job_pool(const ne_job& J, int div, uint64_t* samples) {
  const DevScene& ds = J.scene->dev;
  const uint64_t total = uint64_t(J.width) * uint64_t(J.height) * uint64_t((J.spp + div - 1) / div);
  const uint64_t P = std::min<uint64_t>(default_pool(ds, total), total);
  if (samples) *samples = total;
  return (P + kChunk - 1) / kChunk * kChunk;
}

// One job through the wavefront pipeline on a lane (called from the lane's
// own host thread): the FP32 ne_render path at ceil(spp / div) samples into
// the job's fixed-point batch film, a guided job with its frozen tabulated
// proposal.  A job whose samples fit in kMaxSampleBuffer uses per-sample
// slots, summed into the film at the end (a property of the job alone, so
// its result does not depend on the batch); larger ones take fixed-point
// atomics.  The lane's pool and slots hold the job (sized before the run).
hipError_t lane_render(ne_batch* b, BatchLane& L, size_t j, int div, uint32_t refill_blocks) {
  const ne_job& J = b->jobs[j];
  const DevScene& ds = J.scene->dev;
  uint64_t total = 0;
  const uint64_t P = lane_job_pool(J, div, &total);
  if (P > L.pool_P) return hipErrorInvalidValue;
  hipError_t e;
  WfState w{};
  wf_carve(L.pool, uint32_t(P), w);
  WfArgs a{};
  a.W = J.width; a.H = J.height; a.x0 = 0; a.y0 = 0; a.x1 = J.width; a.y1 = J.height;
  a.max_bounces = J.max_bounces; a.sample_offset = 0; a.job = J.job; a.nchunks = uint32_t(P / kChunk);
  a.seed = J.seed;
  a.sort = 0;
  a.refill_min = kRefillMinIdle;
  a.refill = std::min<uint32_t>(kMaxRefill, std::max<uint32_t>(1, (a.nchunks + refill_blocks - 1) / refill_blocks));
  a.total_paths = total;
  a.counters = nullptr;
  a.film_fixed = b->d_films + b->film_off[j];
  a.film = nullptr;  // the fixed-point sums stay in the batch film
  const size_t sbytes = size_t(total) * 3 * sizeof(float);
  a.sbuf = nullptr;
  if (sbytes <= kMaxSampleBuffer) {
    if (sbytes > L.sbuf_bytes) return hipErrorInvalidValue;
    a.sbuf = static_cast<float*>(L.sbuf);
  }
  a.gstate[0] = a.gstate[1] = nullptr;
  WfGuideHook hook{};
  WfGuideHook* gh = nullptr;
  if (J.guide) {  // frozen: no training records, no steps
    hook.g.net = J.guide->net;
    hook.g.batch = J.guide->batch;
    hook.g.train = 0;
    hook.g.max_depth = J.guide->cfg.guided_vertices;
    hook.g.prob = J.guide->cfg.guide_prob;
    hook.g.cache = J.guide->cache;
    gh = &hook;
  }
  if ((e = launch_wavefront(ds, make_camera(&J.camera), w, a, false, L.stream, L.status, nullptr, gh)) != hipSuccess)
    return e;
  return hipStreamSynchronize(L.stream);
}

// The wavefront executor: NE_BATCH_LANES host threads, one per lane, take the
// jobs longest first (pixels x spp x max bounces) from a shared counter.
ne_status batch_run_lanes(ne_batch* b, int div, float* ms) {
  ne_ctx* c = b->ctx;
  const size_t n = b->jobs.size();
  const int nl = int(std::min<size_t>(NE_BATCH_LANES, n));
  for (int i = int(b->lanes.size()); i < nl; ++i) {
    BatchLane L;
    void* st = nullptr;
    hipError_t e = hipStreamCreate(&L.stream);
    if (e == hipSuccess) e = hipHostMalloc(&st, 64);
    if (e != hipSuccess) {  // a lane is kept only whole
      if (L.stream) (void)hipStreamDestroy(L.stream);
      return fail(c, NE_ERR_HIP, std::string("batch lane set-up: ") + hipGetErrorString(e));
    }
    L.status = static_cast<uint32_t*>(st);
    b->lanes.push_back(L);
  }
  if (c->refill_blocks == 0) {
    int cus = 0, per_cu = 0;
    NE_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    NE_HIP(c, wf_refill_blocks_per_cu(&per_cu));
    c->refill_blocks = uint32_t(std::max(1, cus * std::max(1, per_cu)));
  }
  // every lane can take any job: pools and sample slots for the largest, made
  // before the timed run (slots for the jobs that use them, lane_render)
  uint64_t Pmax = 0;
  size_t sbytes = 0;
  for (const ne_job& J : b->jobs) {
    uint64_t t = 0;
    Pmax = std::max(Pmax, lane_job_pool(J, div, &t));
    const size_t jb = size_t(t) * 3 * sizeof(float);
    if (jb <= kMaxSampleBuffer) sbytes = std::max(sbytes, jb);
  }
  for (int i = 0; i < nl; ++i) {
    BatchLane& L = b->lanes[i];
    if (Pmax > L.pool_P) {
      void* old = L.pool;
      L.pool = nullptr;
      L.pool_P = 0;
      if (old) NE_HIP(c, hipFree(old));
      void* nb = nullptr;
      NE_HIP(c, hipMalloc(&nb, wf_bytes(uint32_t(Pmax))));
      L.pool = nb;
      L.pool_P = uint32_t(Pmax);
    }
    if (sbytes > L.sbuf_bytes) {
      void* old = L.sbuf;
      L.sbuf = nullptr;
      L.sbuf_bytes = 0;
      if (old) NE_HIP(c, hipFree(old));
      void* nb = nullptr;
      NE_HIP(c, hipMalloc(&nb, sbytes));
      L.sbuf = nb;
      L.sbuf_bytes = sbytes;
    }
  }
  ne_status st;
  for (const ne_job& J : b->jobs)
    if (J.guide && (st = guide_refresh_cache(c, const_cast<ne_guide*>(J.guide))) != NE_OK) return st;
  NE_HIP(c, hipMemsetAsync(b->d_films, 0, b->film_words * sizeof(long long), c->stream));
  NE_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<size_t> order(n);
  for (size_t j = 0; j < n; ++j) order[j] = j;
  auto work = [&](size_t j) {
    const ne_job& J = b->jobs[j];
    return double(J.width) * J.height * J.spp * double(J.max_bounces + 1);
  };
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work(x) > work(y); });
  std::atomic<size_t> next{0};
  std::atomic<int> failed{int(hipSuccess)};
  std::vector<float> done(n, 0.0f);
  const uint32_t rb = c->refill_blocks;
  const int device = c->device;
  const auto t0 = std::chrono::steady_clock::now();
  auto worker = [&](int i) {
    hipError_t e = hipSetDevice(device);
    while (e == hipSuccess && failed.load() == int(hipSuccess)) {
      const size_t k = next.fetch_add(1);
      if (k >= n) break;
      e = lane_render(b, b->lanes[i], order[k], div, rb);
      done[order[k]] = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (e != hipSuccess) {
      int expect = int(hipSuccess);
      failed.compare_exchange_strong(expect, int(e));
    }
  };
  std::vector<std::thread> threads;
  for (int i = 0; i < nl; ++i) {
    try {
      threads.emplace_back(worker, i);
    } catch (...) {
      worker(i);  // no thread to spare: this one serves the lane (until the queue is empty)
    }
  }
  for (std::thread& t : threads) t.join();
  const float wall = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (failed.load() != int(hipSuccess)) {
    const hipError_t e = hipError_t(failed.load());
    return fail(c, e == hipErrorOutOfMemory ? NE_ERR_OUT_OF_MEMORY : NE_ERR_HIP,
                std::string("batch wavefront lane: ") + hipGetErrorString(e));
  }
  if (ms) *ms = wall;
  b->job_ms = done;
  return NE_OK;
}

ne_status batch_execute(ne_batch* b, const ne_batch_config& cfg, int div, float* ms) {
  return cfg.executor == NE_BATCH_WAVEFRONT ? batch_run_lanes(b, div, ms) : batch_launch(b, cfg, div, ms);
}

}  // namespace

ne_status ne_batch_create(ne_ctx* c, const ne_job* jobs, int32_t n, ne_batch** out) {
  if (!c || !out) return fail(c, NE_ERR_INVALID_ARG, "ctx/out is NULL");
  *out = nullptr;
  if (n < 1 || n > kMaxBatchJobs || !jobs) return fail(c, NE_ERR_INVALID_ARG, "job count must be in [1, 1024]");
  for (int32_t j = 0; j < n; ++j) {
    const ne_job& J = jobs[j];
    if (!J.scene || J.scene->ctx != c) return fail(c, NE_ERR_INVALID_ARG, "job scene is NULL or from another context");
    if (J.width <= 0 || J.height <= 0 || J.width > 32768 || J.height > 32768)
      return fail(c, NE_ERR_INVALID_ARG, "job image size out of range");
    if (J.spp <= 0) return fail(c, NE_ERR_INVALID_ARG, "job spp must be positive");
    if (J.max_bounces < 0 || J.max_bounces > kMaxVertices) return fail(c, NE_ERR_INVALID_ARG, "job max_bounces out of range");
    if (J.guide && (J.guide->ctx != c || !J.guide->cache_mix))
      return fail(c, NE_ERR_INVALID_ARG, "a job's guide must come from the same context and tabulate its proposal");
  }
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  ne_batch* b = new ne_batch();
  b->ctx = c;
  b->jobs.assign(jobs, jobs + n);
  for (int32_t j = 0; j < n; ++j) {
    b->film_off.push_back(b->film_words);
    b->film_words += size_t(jobs[j].width) * size_t(jobs[j].height) * 3;
  }
  auto bail = [&](ne_status st) { batch_free(b); delete b; return st; };
  b->surfaces = false;
  b->guided = false;
  for (int32_t j = 0; j < n; ++j) {
    if (scene_has_surfaces(jobs[j].scene->dev)) b->surfaces = true;
    if (jobs[j].guide) b->guided = true;
  }
  int cus = 0, per_cu = 0, khz = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
      batch_blocks_per_cu(b->surfaces, b->guided, &per_cu) != hipSuccess ||
      hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || khz <= 0)
    return bail(fail(c, NE_ERR_HIP, "device attribute query failed"));
  b->grid_blocks = uint32_t(std::max(1, cus * std::max(1, per_cu)));
  b->ticks_per_ms = double(khz);
  if (hipMalloc(&b->d_films, b->film_words * sizeof(long long)) != hipSuccess ||
      hipMalloc(&b->d_jobs, size_t(n) * sizeof(BatchJob)) != hipSuccess ||
      hipMalloc(&b->d_misc, 2 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&b->d_job_ms, size_t(n) * sizeof(float)) != hipSuccess)
    return bail(fail(c, NE_ERR_OUT_OF_MEMORY, "batch allocation failed"));
  if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
    return bail(fail(c, NE_ERR_HIP, "event creation failed"));
  c->refs.fetch_add(1);
  *out = b;
  return NE_OK;
}

void ne_batch_release(ne_batch* b) {
  if (!b) return;
  ne_ctx* c = b->ctx;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    batch_free(b);
  }
  delete b;
  ctx_unref(c);
}

ne_status ne_batch_autotune(ne_batch* b, int32_t div, ne_batch_config* best, float* best_ms) {
  if (!b) return fail(nullptr, NE_ERR_INVALID_ARG, "batch is NULL");
  ne_ctx* c = b->ctx;
  if (div < 1) return fail(c, NE_ERR_INVALID_ARG, "spp_divisor must be >= 1");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  // candidates: the persistent kernel with square and wide tiles, few to many
  // samples per item, and the wavefront executor
  static const ne_batch_config cands[] = {
      {8, 8, 1, NE_BATCH_PERSISTENT},   {8, 8, 4, NE_BATCH_PERSISTENT},   {8, 8, 16, NE_BATCH_PERSISTENT},
      {16, 4, 4, NE_BATCH_PERSISTENT},  {16, 4, 16, NE_BATCH_PERSISTENT}, {16, 16, 1, NE_BATCH_PERSISTENT},
      {16, 16, 4, NE_BATCH_PERSISTENT}, {32, 8, 4, NE_BATCH_PERSISTENT},  {8, 8, 4, NE_BATCH_WAVEFRONT}};
  float pms = 0, wms = 0;  // calibration times: fastest persistent shape, wavefront executor
  ne_batch_config pc = cands[0], wc = cands[0];
  bool first = true;
  for (const ne_batch_config& cfg : cands) {
    float ms = 0;
    ne_status st = batch_execute(b, cfg, div, nullptr);  // graph build / pools + warm-up
    if (st == NE_OK) st = batch_execute(b, cfg, div, &ms);
    if (st != NE_OK) return st;
    if (cfg.executor == NE_BATCH_WAVEFRONT) {
      wms = ms;
      wc = cfg;
    } else if (first || ms < pms) {
      pms = ms;
      pc = cfg;
      first = false;
    }
  }
  // The wavefront executor's per-job costs do not shrink with the spp, so a
  // calibration run favours the persistent kernel: the two are compared again
  // on one full-spp run each.
  bool wave = wms < pms;
  if (div > 1) {
    float pf = 0, wf = 0;
    ne_status st = batch_execute(b, pc, 1, &pf);
    if (st == NE_OK) st = batch_execute(b, wc, 1, nullptr);  // pools sized for full spp before timing
    if (st == NE_OK) st = batch_execute(b, wc, 1, &wf);
    if (st != NE_OK) return st;
    wave = wf < pf;
  }
  b->tuned = wave ? wc : pc;
  if (best) *best = b->tuned;
  if (best_ms) *best_ms = wave ? wms : pms;
  return NE_OK;
}

ne_status ne_batch_run(ne_batch* b, const ne_batch_config* cfg, float* ms) {
  if (!b) return fail(nullptr, NE_ERR_INVALID_ARG, "batch is NULL");
  ne_ctx* c = b->ctx;
  const ne_batch_config use = cfg ? *cfg : b->tuned;
  if (!cfg_valid(use)) return fail(c, NE_ERR_INVALID_ARG, "batch item shape out of range");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  return batch_execute(b, use, 1, ms);
}

ne_status ne_batch_job_times(ne_batch* b, float* ms) {
  if (!b || !ms) return fail(b ? b->ctx : nullptr, NE_ERR_INVALID_ARG, "batch/ms is NULL");
  std::lock_guard<std::mutex> lk(b->ctx->mu);
  if (b->job_ms.empty()) return fail(b->ctx, NE_ERR_INVALID_ARG, "the batch has not run");
  std::copy(b->job_ms.begin(), b->job_ms.end(), ms);
  return NE_OK;
}

ne_status ne_batch_read(ne_batch* b, int32_t job, double* sum) {
  if (!b || !sum) return fail(b ? b->ctx : nullptr, NE_ERR_INVALID_ARG, "batch/sum is NULL");
  ne_ctx* c = b->ctx;
  if (job < 0 || size_t(job) >= b->jobs.size()) return fail(c, NE_ERR_INVALID_ARG, "job index out of range");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  const size_t n = size_t(b->jobs[job].width) * size_t(b->jobs[job].height) * 3;
  DevBuf t;
  ne_status st = tmp_out<double>(c, t, n);
  if (st != NE_OK) return st;
  NE_HIP(c, launch_fixed_to_double(b->d_films + b->film_off[job], static_cast<double*>(t.p), n, c->stream));
  if ((st = tmp_back(c, sum, t, n)) != NE_OK) return st;
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

// ---- learned guide -------------------------------------------------------------

namespace {

void guide_free(ne_guide* g) {
  for (void* p : g->allocs) (void)hipFree(p);
  g->allocs.clear();
}

template <class T>
bool guide_alloc(ne_guide* g, T*& p, size_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return false;
  g->allocs.push_back(q);
  p = static_cast<T*>(q);
  return true;
}

}  // namespace

ne_status ne_guide_create(ne_ctx* c, const ne_guide_config* cfg_in, ne_guide** out) {
  if (!c || !out) return fail(c, NE_ERR_INVALID_ARG, "ctx/out is NULL");
  *out = nullptr;
  ne_guide_config cfg{19, 16, 512, 65536, 1e-2f, 1e-3f, 0.9f, 0.99f, 1e-8f, 1, 0.25f, 0, 0, 4};
  if (cfg_in) cfg = *cfg_in;
  if (cfg.cache_res < 0 || cfg.cache_res > 128) return fail(c, NE_ERR_INVALID_ARG, "cache_res must be in [0, 128]");
  if (cfg.cache_res > 0 && (cfg.cache_dirs < 1 || cfg.cache_dirs > 16))
    return fail(c, NE_ERR_INVALID_ARG, "cache_dirs must be in [1, 16]");
  if (cfg.log2_table < 10 || cfg.log2_table > 22) return fail(c, NE_ERR_INVALID_ARG, "log2_table must be in [10, 22]");
  if (cfg.base_res < 2 || cfg.max_res < cfg.base_res || cfg.max_res > 4096)
    return fail(c, NE_ERR_INVALID_ARG, "level resolutions out of range");
  if (cfg.batch < 16 || cfg.batch > (1 << 22)) return fail(c, NE_ERR_INVALID_ARG, "batch must be in [16, 2^22]");
  if (cfg.guided_vertices < 0 || cfg.guided_vertices > kMaxVertices)
    return fail(c, NE_ERR_INVALID_ARG, "guided_vertices out of range");
  if (!(cfg.guide_prob >= 0.0f && cfg.guide_prob <= 0.9f)) return fail(c, NE_ERR_INVALID_ARG, "guide_prob must be in [0, 0.9]");
  if (!(cfg.lr_table >= 0) || !(cfg.lr_mlp >= 0) || !(cfg.beta1 >= 0 && cfg.beta1 < 1) ||
      !(cfg.beta2 >= 0 && cfg.beta2 < 1) || !(cfg.eps > 0))
    return fail(c, NE_ERR_INVALID_ARG, "Adam parameters out of range");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  ne_guide* g = new ne_guide();
  g->ctx = c;
  g->cfg = cfg;
  cfg.batch = (cfg.batch + 15) / 16 * 16;
  g->cfg.batch = cfg.batch;
  GuideNet& n = g->net;
  n.log2T = uint32_t(cfg.log2_table);
  const double growth = std::exp((std::log(double(cfg.max_res)) - std::log(double(cfg.base_res))) / (kGuideLevels - 1));
  for (int l = 0; l < kGuideLevels; ++l) n.res[l] = int32_t(std::floor(double(cfg.base_res) * std::pow(growth, l) + 1e-6));
  g->table_n = size_t(kGuideLevels) * (size_t(1) << n.log2T) * kGuideFeat;
  const size_t cap = size_t(cfg.batch);
  GuideTrain& t = g->train;
  t.waves = 512;
  auto bail = [&](ne_status st) { guide_free(g); delete g; return st; };
  bool ok = guide_alloc(g, n.table, g->table_n) && guide_alloc(g, n.table_bf, g->table_n / kGuideFeat) &&
            guide_alloc(g, n.table_grad, g->table_n) &&
            guide_alloc(g, n.table_m, g->table_n) && guide_alloc(g, n.table_v, g->table_n) &&
            guide_alloc(g, n.mlp, kGuideParams) && guide_alloc(g, n.mlp_grad, kGuideParams) &&
            guide_alloc(g, n.mlp_m, kGuideParams) && guide_alloc(g, n.mlp_v, kGuideParams) &&
            guide_alloc(g, n.wf, kGuideWeights) && guide_alloc(g, n.wb, kGuideWeights) &&
            guide_alloc(g, g->batch.s, 10 * cap) && guide_alloc(g, g->batch.count, 1) &&
            guide_alloc(g, t.act, size_t(kGuideLayers) * cap * kGuideWidth) &&
            guide_alloc(g, t.dy, cap * kGuideWidth) && guide_alloc(g, t.dy2, cap * kGuideWidth) &&
            guide_alloc(g, t.denc, cap * kGuideEnc) && guide_alloc(g, t.partial, size_t(t.waves) * kGuideParams) &&
            guide_alloc(g, t.loss, 2);
  if (ok && cfg.cache_res > 0) {
    const size_t entries = size_t(cfg.cache_res) * cfg.cache_res * cfg.cache_res * size_t(cfg.cache_dirs) * cfg.cache_dirs;
    ok = guide_alloc(g, g->cache_mix, entries * 40);
    g->cache.mix = g->cache_mix;
    g->cache.nc = cfg.cache_res;
    g->cache.nd = cfg.cache_dirs;
  }
  if (!ok) return bail(fail(c, NE_ERR_OUT_OF_MEMORY, "guide allocation failed"));
  g->batch.cap = uint32_t(cap);
  if (hipMemsetAsync(g->batch.count, 0, sizeof(uint32_t), c->stream) != hipSuccess ||
      launch_guide_init(n, cfg.seed, c->stream) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
    return bail(fail(c, NE_ERR_HIP, "guide initialisation failed"));
  c->refs.fetch_add(1);
  *out = g;
  return NE_OK;
}

void ne_guide_release(ne_guide* g) {
  if (!g) return;
  ne_ctx* c = g->ctx;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    guide_free(g);
  }
  delete g;
  ctx_unref(c);
}

ne_status ne_guide_reset(ne_guide* g) {
  if (!g) return fail(nullptr, NE_ERR_INVALID_ARG, "guide is NULL");
  ne_ctx* c = g->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  NE_HIP(c, launch_guide_init(g->net, g->cfg.seed, c->stream));
  NE_HIP(c, hipStreamSynchronize(c->stream));
  g->steps = 0;
  g->samples = 0;
  g->last_loss = 0.0f;
  g->version += 1;
  return NE_OK;
}

ne_status ne_guide_get_stats(ne_guide* g, ne_guide_stats* out) {
  if (!g || !out) return fail(g ? g->ctx : nullptr, NE_ERR_INVALID_ARG, "guide/out is NULL");
  std::lock_guard<std::mutex> lk(g->ctx->mu);
  out->steps = g->steps;
  out->samples = g->samples;
  out->last_loss = g->last_loss;
  out->pad = 0.0f;
  return NE_OK;
}

ne_status ne_guide_eval(ne_guide* g, int32_t n, const float* q, int32_t raw, float* out) {
  if (!g || n < 0 || (n > 0 && (!q || !out))) return fail(g ? g->ctx : nullptr, NE_ERR_INVALID_ARG, "bad guide_eval arguments");
  if (n == 0) return NE_OK;
  ne_ctx* c = g->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  DevBuf bi, bo;
  ne_status st = tmp_in(c, bi, q, size_t(n) * 6);
  if (st == NE_OK) st = tmp_out<float>(c, bo, size_t(n) * kGuideOutputs);
  if (st != NE_OK) return st;
  if (raw) NE_HIP(c, launch_guide_raw(g->net, n, static_cast<float*>(bi.p), static_cast<float*>(bo.p), c->stream));
  else NE_HIP(c, launch_guide_eval(g->net, n, static_cast<float*>(bi.p), static_cast<float*>(bo.p), c->stream));
  if ((st = tmp_back(c, out, bo, size_t(n) * kGuideOutputs)) != NE_OK) return st;
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

ne_status ne_guide_train_step(ne_guide* g, int32_t n, const float* smp, int32_t update, float* loss) {
  if (!g) return fail(nullptr, NE_ERR_INVALID_ARG, "guide is NULL");
  ne_ctx* c = g->ctx;
  if (n < 1 || uint32_t(n) > g->batch.cap || !smp) return fail(c, NE_ERR_INVALID_ARG, "sample count must be in [1, batch]");
  for (int32_t i = 0; i < n; ++i)
    if (!(smp[10 * i + 9] >= 0.0f) || !std::isfinite(smp[10 * i + 9]))
      return fail(c, NE_ERR_INVALID_ARG, "sample weights must be finite and >= 0");
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  // samples: AoS from the caller -> SoA batch buffer
  const size_t cap = g->batch.cap;
  std::vector<float> soa(10 * cap, 0.0f);
  for (int32_t i = 0; i < n; ++i)
    for (int k = 0; k < 10; ++k) soa[size_t(k) * cap + size_t(i)] = smp[10 * i + k];
  const uint32_t cnt = uint32_t(n);
  NE_HIP(c, hipMemcpyAsync(g->batch.s, soa.data(), soa.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
  NE_HIP(c, hipMemcpyAsync(g->batch.count, &cnt, sizeof(cnt), hipMemcpyHostToDevice, c->stream));
  NE_HIP(c, hipMemsetAsync(g->net.table_grad, 0, g->table_n * sizeof(float), c->stream));
  NE_HIP(c, launch_guide_grad(g->net, g->batch, g->train, c->stream));
  float lw[2] = {0, 0};
  NE_HIP(c, hipMemcpyAsync(lw, g->train.loss, sizeof(lw), hipMemcpyDeviceToHost, c->stream));
  NE_HIP(c, hipStreamSynchronize(c->stream));
  g->last_loss = lw[1] > 0 ? lw[0] / lw[1] : 0.0f;
  if (loss) *loss = g->last_loss;
  g->samples += uint64_t(n);
  if (update) {
    g->steps += 1;
    g->version += 1;
    NE_HIP(c, launch_guide_adam(g->net, g->cfg.lr_table, g->cfg.lr_mlp, g->cfg.beta1, g->cfg.beta2, g->cfg.eps,
                                int(g->steps), c->stream));
    NE_HIP(c, hipStreamSynchronize(c->stream));
  }
  return NE_OK;
}

ne_status ne_guide_sizes(ne_guide* g, uint64_t* table_floats, int32_t* mlp_floats, int32_t* level_res) {
  if (!g) return fail(nullptr, NE_ERR_INVALID_ARG, "guide is NULL");
  if (table_floats) *table_floats = g->table_n;
  if (mlp_floats) *mlp_floats = kGuideParams;
  if (level_res)
    for (int l = 0; l < kGuideLevels; ++l) level_res[l] = g->net.res[l];
  return NE_OK;
}

ne_status ne_guide_get_params(ne_guide* g, float* table, float* mlp) {
  if (!g) return fail(nullptr, NE_ERR_INVALID_ARG, "guide is NULL");
  ne_ctx* c = g->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  if (table) NE_HIP(c, hipMemcpy(table, g->net.table, g->table_n * sizeof(float), hipMemcpyDeviceToHost));
  if (mlp) NE_HIP(c, hipMemcpy(mlp, g->net.mlp, kGuideParams * sizeof(float), hipMemcpyDeviceToHost));
  return NE_OK;
}

ne_status ne_guide_set_params(ne_guide* g, const float* table, const float* mlp) {
  if (!g) return fail(nullptr, NE_ERR_INVALID_ARG, "guide is NULL");
  ne_ctx* c = g->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  if (table) NE_HIP(c, hipMemcpy(g->net.table, table, g->table_n * sizeof(float), hipMemcpyHostToDevice));
  if (mlp) NE_HIP(c, hipMemcpy(g->net.mlp, mlp, kGuideParams * sizeof(float), hipMemcpyHostToDevice));
  NE_HIP(c, launch_guide_pack(g->net, c->stream));
  NE_HIP(c, hipStreamSynchronize(c->stream));
  g->version += 1;
  return NE_OK;
}

ne_status ne_guide_get_grads(ne_guide* g, float* table_grad, float* mlp_grad) {
  if (!g) return fail(nullptr, NE_ERR_INVALID_ARG, "guide is NULL");
  ne_ctx* c = g->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  if (table_grad) NE_HIP(c, hipMemcpy(table_grad, g->net.table_grad, g->table_n * sizeof(float), hipMemcpyDeviceToHost));
  if (mlp_grad) NE_HIP(c, hipMemcpy(mlp_grad, g->net.mlp_grad, kGuideParams * sizeof(float), hipMemcpyDeviceToHost));
  return NE_OK;
}

// Re-tabulates the guide's proposal if its parameters changed since the last
// tabulation (context lock held).
static ne_status guide_refresh_cache(ne_ctx* c, ne_guide* g) {
  if (!g->cache_mix || g->cache_version == g->version) return NE_OK;
  NE_HIP(c, launch_guide_bake(g->net, g->cache, g->cache_mix, c->stream));
  g->cache_version = g->version;
  return NE_OK;
}

static GuideUse guide_use(const ne_guide* g) {
  GuideUse u{};
  u.cache = g->cache;
  u.max_depth = g->cfg.guided_vertices;
  u.prob = g->cfg.guide_prob;
  return u;
}

ne_status ne_guide_get_cache(ne_guide* g, float* mix, uint64_t* entries) {
  if (!g) return fail(nullptr, NE_ERR_INVALID_ARG, "guide is NULL");
  ne_ctx* c = g->ctx;
  if (!g->cache_mix) return fail(c, NE_ERR_INVALID_ARG, "the guide tabulates no proposal (cache_res = 0)");
  const uint64_t n = uint64_t(g->cache.nc) * g->cache.nc * g->cache.nc * uint64_t(g->cache.nd) * g->cache.nd;
  if (entries) *entries = n;
  if (!mix) return NE_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  NE_HIP(c, hipSetDevice(c->device));
  ne_status st = guide_refresh_cache(c, g);
  if (st != NE_OK) return st;
  NE_HIP(c, hipMemcpyAsync(mix, g->cache_mix, size_t(n) * 40 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  NE_HIP(c, hipStreamSynchronize(c->stream));
  return NE_OK;
}

ne_status ne_render_guided(ne_ctx* c, const ne_scene* s, const ne_camera_desc* cam, const ne_render_params* p,
                           ne_film* f, ne_guide* g, int32_t train) {
  if (!c) return fail(nullptr, NE_ERR_INVALID_ARG, "ctx is NULL");
  if (!g || g->ctx != c) return fail(c, NE_ERR_INVALID_ARG, "guide is NULL or from another context");
  if (!p || p->precision != NE_FP32) return fail(c, NE_ERR_INVALID_ARG, "guided rendering is FP32");
  if (p->mode == NE_MODE_MEGAKERNEL && (!g->cache_mix || train))
    return fail(c, NE_ERR_INVALID_ARG,
                "the guided megakernel needs a tabulated proposal (cache_res > 0) and does not train");
  std::lock_guard<std::mutex> lk(c->mu);
  ne_status st = check_render(c, s, cam, p, f);
  if (st != NE_OK) return st;
  NE_HIP(c, hipSetDevice(c->device));
  if ((st = guide_refresh_cache(c, g)) != NE_OK) return st;
  if (p->mode == NE_MODE_MEGAKERNEL) {
    const GuideUse u = guide_use(g);
    return render_locked(c, s, cam, p, f, &u);
  }
  const bool count = (p->flags & NE_FLAG_COUNT) != 0;
  const bool timed = (p->flags & NE_FLAG_TIME) != 0;
  if (count) {
    NE_HIP(c, hipMemsetAsync(c->d_counters, 0, sizeof(DevCounters), c->stream));
    c->counters_valid = true;
  }
  StageTimer timer{&c->ev_pool, &c->ev_stages, c->stream};
  StageTimer* tm = nullptr;
  c->stage_timed = false;
  if (p->flags & NE_FLAG_STAGE_TIME) {
    c->ev_stages.clear();
    tm = &timer;
  }
  WfGuideHook hook{};
  hook.g.net = g->net;
  hook.g.batch = g->batch;
  hook.g.train = train ? 1 : 0;
  hook.g.max_depth = g->cfg.guided_vertices;
  hook.g.prob = g->cfg.guide_prob;
  hook.g.cache = g->cache;
  hook.cache_mix = g->cache_mix;
  hook.train = g->train;
  hook.lr_table = g->cfg.lr_table; hook.lr_mlp = g->cfg.lr_mlp;
  hook.beta1 = g->cfg.beta1; hook.beta2 = g->cfg.beta2; hook.eps = g->cfg.eps;
  hook.min_batch = g->batch.cap / 2;
  hook.steps = g->steps;
  hook.samples = 0;
  if (train) NE_HIP(c, hipMemsetAsync(g->batch.count, 0, sizeof(uint32_t), c->stream));
  const DevCamera dc = make_camera(cam);
  const uint64_t steps0 = g->steps;
  st = render_wavefront(c, s, dc, p, f, count, timed, tm, &hook);
  c->stage_timed = st == NE_OK && tm;
  g->steps = hook.steps;
  g->samples += hook.samples;
  if (g->steps > steps0) {
    // every step of the render re-tabulated the proposal right after its update
    g->version += g->steps - steps0;
    if (g->cache_mix && st == NE_OK) g->cache_version = g->version;
  }
  if (st == NE_OK && g->steps > steps0) {  // loss of the last step taken during the render
    float lw[2] = {0, 0};
    NE_HIP(c, hipMemcpyAsync(lw, g->train.loss, sizeof(lw), hipMemcpyDeviceToHost, c->stream));
    NE_HIP(c, hipStreamSynchronize(c->stream));
    g->last_loss = lw[1] > 0 ? lw[0] / lw[1] : 0.0f;
  }
  return st;
}
