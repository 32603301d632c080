// oracle_capi.cpp -- extern "C" surface of the oracle (ne_oracle.hpp) for
// the Python test-suite (ctypes).  TEST INFRASTRUCTURE ONLY: the product
// library never links this file.
#include <cstdint>
#include <memory>

#include "ne_oracle.hpp"

using namespace neo;

static inline V3f v3(const float* p) { return V3f(p[0], p[1], p[2]); }
static inline void put(float* p, V3f v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

extern "C" {

int neo_sizeof_scene_desc() { return int(sizeof(SceneDesc)); }
int neo_sizeof_camera_desc() { return int(sizeof(CameraDesc)); }

void neo_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox4x32_10(ctr, key, out); }

void neo_rng_uniforms(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t job, int n, float* out) {
  Rng r(seed, pixel, sample, job);
  for (int i = 0; i < n; ++i) out[i] = r.next();
}

// ---- per-primitive restatements (rays: o xyz, d xyz) -----------------------
void neo_aabb_isect(int n, const float* ray, const float* box, int* hit, float* tn, float* tf, float* pt) {
  for (int i = 0; i < n; ++i) {
    Hit h = ref_aabb_intersect(v3(box + 6 * i), v3(box + 6 * i + 3), v3(ray + 6 * i), v3(ray + 6 * i + 3));
    hit[i] = h.hit; tn[i] = h.tNear; tf[i] = h.tFar;
    if (pt) put(pt + 3 * i, h.point);
  }
}
void neo_tri_isect(int n, const float* ray, const float* tri, int* hit, float* t, float* nrm) {
  for (int i = 0; i < n; ++i) {
    const float* q = tri + 9 * i;
    Hit h = ref_triangle_intersect(v3(q), v3(q + 3), v3(q + 6), v3(ray + 6 * i), v3(ray + 6 * i + 3));
    hit[i] = h.hit; t[i] = h.hit ? h.tNear : kINF;
    if (nrm) put(nrm + 3 * i, h.normal);
  }
}
void neo_sphere_isect(int n, const float* ray, const float* sph, int* hit, float* tn, float* tf) {
  for (int i = 0; i < n; ++i) {
    Hit h = ref_sphere_intersect(v3(sph + 4 * i), sph[4 * i + 3], v3(ray + 6 * i), v3(ray + 6 * i + 3));
    hit[i] = h.hit; tn[i] = h.tNear; tf[i] = h.tFar;
  }
}
void neo_rect_isect(int n, const float* ray, const float* rect, int* hit, float* t) {
  for (int i = 0; i < n; ++i) {
    const float* q = rect + 9 * i;
    Hit h = ref_rect_intersect(v3(q), v3(q + 3), v3(q + 6), v3(ray + 6 * i), v3(ray + 6 * i + 3));
    hit[i] = h.hit; t[i] = h.hit ? h.tNear : kINF;
  }
}
void neo_triangle_area(int n, const float* tri, float* out) {
  for (int i = 0; i < n; ++i) out[i] = ref_triangle_area(v3(tri + 9 * i), v3(tri + 9 * i + 3), v3(tri + 9 * i + 6));
}
void neo_barycentric(int n, const float* p, const float* tri, float* out) {
  for (int i = 0; i < n; ++i)
    put(out + 3 * i, ref_barycentric(v3(p + 3 * i), v3(tri + 9 * i), v3(tri + 9 * i + 3), v3(tri + 9 * i + 6)));
}
void neo_triangle_contains(int n, const float* p, const float* tri, int* out) {
  for (int i = 0; i < n; ++i)
    out[i] = ref_triangle_contains(v3(tri + 9 * i), v3(tri + 9 * i + 3), v3(tri + 9 * i + 6), v3(p + 3 * i));
}
void neo_onb(int n, const float* nrm, float* v, float* u) {
  for (int i = 0; i < n; ++i) {
    V3f a, b;
    ref_onb(v3(nrm + 3 * i), a, b);
    put(v + 3 * i, a); put(u + 3 * i, b);
  }
}
void neo_to_lcs(int n, const float* v, const float* ns, const float* ss, const float* ts, float* out) {
  for (int i = 0; i < n; ++i) put(out + 3 * i, to_lcs(v3(v + 3 * i), v3(ns + 3 * i), v3(ss + 3 * i), v3(ts + 3 * i)));
}
void neo_to_world(int n, const float* v, const float* ns, const float* ss, const float* ts, float* out) {
  for (int i = 0; i < n; ++i) put(out + 3 * i, to_world(v3(v + 3 * i), v3(ns + 3 * i), v3(ss + 3 * i), v3(ts + 3 * i)));
}
void neo_sample_unit_sphere(int n, const float* e, float* out) {
  for (int i = 0; i < n; ++i) put(out + 3 * i, ref_sample_unit_sphere(e[2 * i], e[2 * i + 1]));
}
void neo_hg_eval(int n, float g, const float* a, const float* b, float* out) {
  for (int i = 0; i < n; ++i) out[i] = ref_hg_eval(g, v3(a + 3 * i), v3(b + 3 * i));
}
void neo_hg_sample_local(int n, float g, const float* u, float* out) {
  for (int i = 0; i < n; ++i) put(out + 3 * i, ref_hg_sample_local(g, u[2 * i], u[2 * i + 1]));
}
void neo_ggx_D(int n, const float* alpha, const float* h, float* out) {
  for (int i = 0; i < n; ++i) out[i] = ref_ggx_D(alpha[i], v3(h + 3 * i));
}
void neo_ggx_G(int n, const float* alpha, const float* wi, const float* wo, float* out) {
  for (int i = 0; i < n; ++i) out[i] = ref_ggx_G(alpha[i], v3(wi + 3 * i), v3(wo + 3 * i));
}
void neo_ggx_pdf(int n, const float* alpha, const float* sc, const float* h, float* out) {
  for (int i = 0; i < n; ++i) out[i] = ref_ggx_pdf(alpha[i], v3(sc + 3 * i), v3(h + 3 * i));
}
void neo_ggx_sample_h(int n, const float* alpha, const float* u, float* out) {
  for (int i = 0; i < n; ++i) put(out + 3 * i, ref_ggx_sample_h(alpha[i], u[2 * i], u[2 * i + 1]));
}
void neo_glossy_eval(int n, const float* rough, const float* metal, const float* albedo, const float* wi,
                     const float* wo, float* out) {
  for (int i = 0; i < n; ++i)
    put(out + 3 * i, ref_glossy_eval(rough[i], metal[i], v3(albedo + 3 * i), v3(wi + 3 * i), v3(wo + 3 * i)));
}
void neo_interp_density(int W, int H, int D, const float* grid, int n, const float* gcs, float* out) {
  Grid g; g.W = W; g.H = H; g.D = D; g.data = grid;
  for (int i = 0; i < n; ++i) out[i] = ref_interpolated_density(g, v3(gcs + 3 * i));
}
float neo_max_density(int W, int H, int D, const float* grid) {
  Grid g; g.W = W; g.H = H; g.D = D; g.data = grid;
  return ref_max_density(g);
}
// Reference delta tracking on OCS rays; one Philox stream per ray.
void neo_delta_track(int W, int H, int D, const float* grid, float inv_max, float sigma_bar, int n, const float* ray,
                     const float* tint, uint64_t seed, int* scattered, float* t_out) {
  Grid g; g.W = W; g.H = H; g.D = D; g.data = grid;
  for (int i = 0; i < n; ++i) {
    Rng r(seed, uint32_t(i), 0, 0);
    float t;
    scattered[i] = ref_delta_track(g, inv_max, sigma_bar, v3(ray + 6 * i), v3(ray + 6 * i + 3), tint[2 * i],
                                   tint[2 * i + 1], r, &t);
    t_out[i] = t;
  }
}
void neo_ratio_track(int W, int H, int D, const float* grid, float inv_max, float sigma_bar, int n, const float* ray,
                     const float* tint, uint64_t seed, float* tr) {
  Grid g; g.W = W; g.H = H; g.D = D; g.data = grid;
  for (int i = 0; i < n; ++i) {
    Rng r(seed, uint32_t(i), 0, 0);
    tr[i] = ref_ratio_track(g, inv_max, sigma_bar, v3(ray + 6 * i), v3(ray + 6 * i + 3), tint[2 * i], tint[2 * i + 1], r);
  }
}
void neo_homogeneous_tr(int n, const float* ext, const float* density, const float* dist, float* out) {
  for (int i = 0; i < n; ++i) put(out + 3 * i, ref_homogeneous_tr(v3(ext + 3 * i), density[i], dist[i]));
}
void neo_homogeneous_sample(int n, const float* sig_s, const float* sig_a, float density, const float* tint,
                            uint64_t seed, int* sampled, float* t_out, float* weight) {
  for (int i = 0; i < n; ++i) {
    Rng r(seed, uint32_t(i), 0, 0);
    V3f w;
    sampled[i] = ref_homogeneous_sample(v3(sig_s), v3(sig_a), density, tint[2 * i], tint[2 * i + 1], r, t_out + i, &w);
    put(weight + 3 * i, w);
  }
}
void neo_post_processing(int n, const float* in, float* out) {
  for (int i = 0; i < n; ++i) put(out + 3 * i, ref_post_processing(v3(in + 3 * i)));
}
void neo_camera_dirs(const CameraDesc* cd, int n, const float* uv, float* out) {
  Camera c = ref_camera(v3(cd->position), v3(cd->look_at), v3(cd->up), cd->vfov, cd->aspect, 1.0f);
  for (int i = 0; i < n; ++i) put(out + 3 * i, ref_camera_dir(c, uv[2 * i], uv[2 * i + 1]));
}
void neo_tile_rect(int W, int H, int index, int* out4) {
  TileRect r = ref_tile_rect(W, H, index);
  out4[0] = r.x0; out4[1] = r.y0; out4[2] = r.x1; out4[3] = r.y1;
}
// corrected-mode BSDF/phase (used by the GPU parity tests)
void neo_hg_p(int n, const float* g, const float* cos_t, float* out) {
  for (int i = 0; i < n; ++i) out[i] = hg_p<float>(g[i], cos_t[i]);
}
void neo_hg_sample(int n, const float* g, const float* d, const float* u, float* out) {
  for (int i = 0; i < n; ++i) put(out + 3 * i, hg_sample<float>(g[i], v3(d + 3 * i), u[2 * i], u[2 * i + 1]));
}
// mat: albedo xyz, roughness, metallic (5 floats); wo, wi local frame
void neo_glossy_c(int n, const float* mat, const float* wo, const float* wi, const float* u, float* f, float* pdf,
                  float* smp) {
  for (int i = 0; i < n; ++i) {
    Glossy<float> m;
    m.albedo = v3(mat + 5 * i);
    m.alpha = mat[5 * i + 3] * mat[5 * i + 3];
    m.metallic = mat[5 * i + 4];
    put(f + 3 * i, m.eval(v3(wo + 3 * i), v3(wi + 3 * i)));
    pdf[i] = m.pdf(v3(wo + 3 * i), v3(wi + 3 * i));
    put(smp + 3 * i, m.sample(v3(wo + 3 * i), u[3 * i], u[3 * i + 1], u[3 * i + 2]));
  }
}

// ---- world (scene) level ---------------------------------------------------
struct OWorld {
  int prec;
  std::unique_ptr<World<float>> wf;
  std::unique_ptr<World<double>> wd;
};

void* neo_world_create(const SceneDesc* s, int fp64) {
  OWorld* w = new OWorld();
  w->prec = fp64;
  if (fp64) w->wd.reset(new World<double>(s));
  else w->wf.reset(new World<float>(s));
  return w;
}
void neo_world_destroy(void* p) { delete static_cast<OWorld*>(p); }

int neo_world_bvh_nodes(void* p) {
  OWorld* w = static_cast<OWorld*>(p);
  return w->prec ? int(w->wd->nodes.size()) : int(w->wf->nodes.size());
}
void neo_world_majorants(void* p, float* out) {
  OWorld* w = static_cast<OWorld*>(p);
  const std::vector<float>& m = w->prec ? w->wd->maj : w->wf->maj;
  for (size_t i = 0; i < m.size(); ++i) out[i] = m[i];
}

// environment: sampled directions (u2: 2 uniforms per sample), solid-angle pdfs and radiance (FP32 world)
void neo_world_env_sample(void* p, int n, const float* u2, float* dirs, float* pdf, float* le) {
  const World<float>& W = *static_cast<OWorld*>(p)->wf;
  for (int i = 0; i < n; ++i) {
    V3f wi, L;
    float pl;
    W.env_sample(u2[2 * i], u2[2 * i + 1], wi, pl, L);
    put(dirs + 3 * i, wi); pdf[i] = pl; put(le + 3 * i, L);
  }
}
void neo_world_env_eval(void* p, int n, const float* dirs, float* pdf, float* le) {
  const World<float>& W = *static_cast<OWorld*>(p)->wf;
  for (int i = 0; i < n; ++i) {
    pdf[i] = W.env_pdf(v3(dirs + 3 * i));
    put(le + 3 * i, W.env_le(v3(dirs + 3 * i)));
  }
}
// the reference's Distribution2D over f (width x height), conditionals of cond_n values
void neo_ref_dist2d(const float* f, int width, int height, int cond_n, int n, const float* u2, float* uv, float* pdf) {
  RefDistribution2D d(f, width, height, cond_n);
  for (int i = 0; i < n; ++i) d.sampleContinuous(u2[2 * i], u2[2 * i + 1], uv + 2 * i, pdf + i);
}

// closest surface hit in (tmin, tmax); type -1 on miss
void neo_world_closest(void* p, int n, const float* ray, float tmin, float tmax, float* t, int* type, int* index) {
  OWorld* w = static_cast<OWorld*>(p);
  for (int i = 0; i < n; ++i) {
    World<float>::SHit h;
    V3f o = v3(ray + 6 * i), d = v3(ray + 6 * i + 3);
    bool ok = w->wf->closest(o, d, tmin, tmax, h);
    t[i] = ok ? h.t : kINF; type[i] = ok ? h.type : -1; index[i] = ok ? int(h.index) : -1;
  }
}
// brute-force closest hit (linear scan) -- independent of the BVH
void neo_world_closest_brute(void* p, int n, const float* ray, float tmin, float tmax, float* t, int* type,
                             int* index) {
  OWorld* w = static_cast<OWorld*>(p);
  const World<float>& W = *w->wf;
  for (int i = 0; i < n; ++i) {
    V3f o = v3(ray + 6 * i), d = v3(ray + 6 * i + 3);
    float best = tmax; int bt = -1, bi = -1; bool found = false;
    uint64_t best_key = ~0ull;
    for (const PrimRef& pr : W.prims) {
      float tt;
      if (W.isect_prim(pr, o, d, tmin, tmax, tt)) {
        uint64_t key = (uint64_t(pr.type) << 32) | pr.index;
        if (tt < best || (found && tt == best && key < best_key)) {
          best = tt; best_key = key; bt = int(pr.type); bi = int(pr.index); found = true;
        }
      }
    }
    t[i] = found ? best : kINF; type[i] = bt; index[i] = bi;
  }
}
// transmittance along rays (surface occlusion x ratio tracking); one Philox stream per ray
void neo_world_transmittance(void* p, int n, const float* ray, const float* tmax, uint64_t seed, double* out) {
  OWorld* w = static_cast<OWorld*>(p);
  for (int i = 0; i < n; ++i) {
    Rng r(seed, uint32_t(i), 0, 0);
    if (w->prec)
      out[i] = w->wd->transmittance(V3<double>(v3(ray + 6 * i)), V3<double>(v3(ray + 6 * i + 3)), double(tmax[i]), r);
    else
      out[i] = w->wf->transmittance(v3(ray + 6 * i), v3(ray + 6 * i + 3), tmax[i], r);
  }
}
// delta-tracking collision distances along rays through the volume (per-cell majorants)
void neo_world_delta_track(void* p, int n, const float* ray, uint64_t seed, int* scattered, double* t_out) {
  OWorld* w = static_cast<OWorld*>(p);
  for (int i = 0; i < n; ++i) {
    Rng r(seed, uint32_t(i), 0, 0);
    scattered[i] = 0; t_out[i] = 0;
    if (w->prec) {
      V3<double> o(v3(ray + 6 * i)), d(v3(ray + 6 * i + 3));
      double ta, tb, th;
      if (w->wd->box_interval(o, d, ta, tb)) {
        ta = std::max(ta, 0.0);
        if (ta < tb && w->wd->delta_track(o, d, ta, tb, r, th)) { scattered[i] = 1; t_out[i] = th; }
      }
    } else {
      V3f o = v3(ray + 6 * i), d = v3(ray + 6 * i + 3);
      float ta, tb, th;
      if (w->wf->box_interval(o, d, ta, tb)) {
        ta = std::max(ta, 0.0f);
        if (ta < tb && w->wf->delta_track(o, d, ta, tb, r, th)) { scattered[i] = 1; t_out[i] = th; }
      }
    }
  }
}
// Integrator::Li for a batch of rays (src/integrators/Integrator.h:23): ray i
// draws from the stream (seed; i, sample, 0), as the device's ne_li keys it.
void neo_li(void* p, int n, const float* ray, int max_bounces, uint64_t seed, uint32_t sample, double* out) {
  OWorld* w = static_cast<OWorld*>(p);
  for (int i = 0; i < n; ++i) {
    Rng r(seed, uint32_t(i), sample, 0);
    if (w->prec) {
      V3<double> L = Li<double>(*w->wd, V3<double>(v3(ray + 6 * i)), V3<double>(v3(ray + 6 * i + 3)), max_bounces, r);
      out[3 * i] = L.x; out[3 * i + 1] = L.y; out[3 * i + 2] = L.z;
    } else {
      V3f L = Li<float>(*w->wf, v3(ray + 6 * i), v3(ray + 6 * i + 3), max_bounces, r);
      out[3 * i] = L.x; out[3 * i + 1] = L.y; out[3 * i + 2] = L.z;
    }
  }
}
// Full render of pixels [x0,x1)x[y0,y1): sum += per-pixel radiance sums.
// stats (optional): [extension rays, shadow rays]
void neo_render(void* p, const CameraDesc* cam, int W, int H, int x0, int y0, int x1, int y1, int spp,
                int max_bounces, uint64_t seed, uint32_t sample_offset, uint32_t job, int nthreads, double* sum,
                uint64_t* stats) {
  OWorld* w = static_cast<OWorld*>(p);
  PathStats ps;
  if (w->prec)
    render<double>(*w->wd, *cam, W, H, x0, y0, x1, y1, spp, max_bounces, seed, sample_offset, job, nthreads, sum,
                   stats ? &ps : nullptr);
  else
    render<float>(*w->wf, *cam, W, H, x0, y0, x1, y1, spp, max_bounces, seed, sample_offset, job, nthreads, sum,
                  stats ? &ps : nullptr);
  if (stats) { stats[0] = ps.ext_rays; stats[1] = ps.shadow_rays; }
}

// Probe pixels at high spp: for each of the n pixels (xy[2k], xy[2k+1]) the
// spp camera samples of render() (same streams, same jitter), summed in FP64
// with the sum of squares of each sample's radiance (for a per-pixel variance).
// Work is spread over threads in chunks of 1024 samples.
void neo_render_pixels(void* p, const CameraDesc* cd, int W, int H, int n, const int* xy, int spp, int max_bounces,
                       uint64_t seed, uint32_t sample_offset, uint32_t job, int nthreads, double* sum, double* sumsq) {
  OWorld* w = static_cast<OWorld*>(p);
  const int chunk = 1024, nch = (spp + chunk - 1) / chunk;
  const size_t items = size_t(n) * size_t(nch);
  std::vector<double> part(items * 6, 0.0);
  V3f pos(cd->position[0], cd->position[1], cd->position[2]);
  V3f at(cd->look_at[0], cd->look_at[1], cd->look_at[2]);
  V3f up(cd->up[0], cd->up[1], cd->up[2]);
  Camera cam = ref_camera(pos, at, up, cd->vfov, cd->aspect, 1.0f);
  std::atomic<size_t> next{0};
  auto worker = [&]() {
    while (true) {
      size_t it = next.fetch_add(1);
      if (it >= items) break;
      const int k = int(it / size_t(nch)), c = int(it % size_t(nch));
      const int x = xy[2 * k], y = xy[2 * k + 1];
      const uint32_t pix = uint32_t(y) * uint32_t(W) + uint32_t(x);
      double* acc = &part[it * 6];
      for (int si = c * chunk; si < std::min(spp, (c + 1) * chunk); ++si) {
        Rng rng(seed, pix, sample_offset + uint32_t(si), job);
        float ju = rng.next(), jv = rng.next();
        V3f dir = ref_camera_dir(cam, (float(x) + ju) / float(W), (float(y) + jv) / float(H));
        double L[3];
        if (w->prec) {
          V3<double> r = Li<double>(*w->wd, V3<double>(pos), V3<double>(dir), max_bounces, rng);
          L[0] = r.x; L[1] = r.y; L[2] = r.z;
        } else {
          V3f r = Li<float>(*w->wf, pos, dir, max_bounces, rng);
          L[0] = r.x; L[1] = r.y; L[2] = r.z;
        }
        for (int a = 0; a < 3; ++a) { acc[a] += L[a]; acc[3 + a] += L[a] * L[a]; }
      }
    }
  };
  const int nt = std::max(1, nthreads);
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
  for (int k = 0; k < n; ++k)
    for (int a = 0; a < 3; ++a) {
      double s1 = 0.0, s2 = 0.0;
      for (int c = 0; c < nch; ++c) {  // fixed order: the result does not depend on thread timing
        s1 += part[(size_t(k) * nch + c) * 6 + a];
        s2 += part[(size_t(k) * nch + c) * 6 + 3 + a];
      }
      sum[3 * k + a] = s1;
      sumsq[3 * k + a] = s2;
    }
}

}  // extern "C"
