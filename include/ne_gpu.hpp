// ne_gpu.hpp -- header-only C++ facade over the C ABI (ne_gpu.h) with the
// reference's renderer interfaces:
//
//   narval_gpu::GpuPathIntegrator   Integrator (src/integrators/Integrator.h:13-25):
//                                   Li(Ray) -> radiance of one path; clone()
//   narval_gpu::OfflineEngine       OfflineEngine (src/core/OfflineEngine.h:35-59):
//                                   constructed from (camera, settings, scene),
//                                   renderTile(camera, index, finished) writes the
//                                   post-processed tile into `pixels` and sets
//                                   `finished`; coreLoop() renders all 400 tiles
//   narval_gpu::SceneReader         SceneReader (src/io/SceneReader.h:8-40):
//                                   loadScene(path, absolutePath), getScene(),
//                                   getMainCamera(), getSettings()
//
// The facade owns C handles through RAII wrappers and turns error statuses
// into narval_gpu::Error exceptions on the C++ side only (the C ABI itself
// never throws).  Vectors are plain float triples so the header needs no glm;
// a maintainer converts glm::vec3 at the call site (see the binding notes).
#ifndef NE_GPU_HPP
#define NE_GPU_HPP

#include <atomic>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ne_gpu.h"

namespace narval_gpu {

struct Vec3 {
  float x = 0, y = 0, z = 0;
};
// src/primitives/Ray.h:17-32
struct Ray {
  Vec3 o, d;
};
// src/core/Settings.h:9-14
struct SceneSettings {
  int width = 512, height = 512;
  int spp = 1;
  int bounces = 1;
  bool hdr = false;
};
// Camera(lookFrom, lookAt, up, vfov, aspect) (src/core/Camera.cpp:7-26)
struct CameraSpec {
  Vec3 position, look_at, up{0, 1, 0};
  float vfov = 45.0f;
  float aspect = 1.0f;
  ne_camera_desc desc() const {
    ne_camera_desc d{};
    d.position[0] = position.x; d.position[1] = position.y; d.position[2] = position.z;
    d.look_at[0] = look_at.x; d.look_at[1] = look_at.y; d.look_at[2] = look_at.z;
    d.up[0] = up.x; d.up[1] = up.y; d.up[2] = up.z;
    d.vfov = vfov;
    d.aspect = aspect;
    return d;
  }
};

class Error : public std::runtime_error {
 public:
  Error(ne_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
  ne_status status;
};

namespace detail {
inline void check(const ne_ctx* ctx, ne_status s) {
  if (s != NE_OK) throw Error(s, ne_last_error(ctx));
}
struct CtxDel { void operator()(ne_ctx* c) const { ne_ctx_destroy(c); } };
struct SceneDel { void operator()(ne_scene* s) const { ne_scene_release(s); } };
struct FilmDel { void operator()(ne_film* f) const { ne_film_release(f); } };
struct FileDel { void operator()(ne_scene_file* f) const { ne_scene_file_release(f); } };
}  // namespace detail

// SceneReader over the library's scene-file reader.  A relative filePath, like
// every asset path inside the file, is taken relative to resourcesDir (the
// reference's RESOURCES_DIR); getScene() is the flattened descriptor that
// Scene uploads, valid while the reader lives.
class SceneReader {
 public:
  explicit SceneReader(std::string resourcesDir = "", int flags = 0)
      : dir_(std::move(resourcesDir)), flags_(flags) {}
  SceneReader(const std::string& filePath, bool absolutePath, std::string resourcesDir = "", int flags = 0)
      : dir_(std::move(resourcesDir)), flags_(flags) {
    loadScene(filePath, absolutePath);
  }
  void loadScene(const std::string& filePath, bool absolutePath = false) {
    std::string path = filePath;
    if (!absolutePath && !dir_.empty() && !filePath.empty() && filePath[0] != '/')
      path = dir_ + (dir_.back() == '/' ? "" : "/") + filePath;
    ne_load_options o{};
    o.flags = flags_;
    o.resources_dir = dir_.empty() ? nullptr : dir_.c_str();
    ne_scene_file* f = nullptr;
    detail::check(nullptr, ne_scene_file_read(path.c_str(), &o, &f));
    file_.reset(f, detail::FileDel());
    detail::check(nullptr, ne_scene_file_get(f, &desc_, &cam_, &set_));
  }
  const ne_scene_desc& getScene() const { return desc_; }
  CameraSpec getMainCamera() const {
    CameraSpec c;
    c.position = {cam_.position[0], cam_.position[1], cam_.position[2]};
    c.look_at = {cam_.look_at[0], cam_.look_at[1], cam_.look_at[2]};
    c.up = {cam_.up[0], cam_.up[1], cam_.up[2]};
    c.vfov = cam_.vfov;
    c.aspect = cam_.aspect;
    return c;
  }
  SceneSettings getSettings() const {
    SceneSettings s;
    s.width = set_.width;
    s.height = set_.height;
    s.spp = set_.spp;
    s.bounces = set_.bounces;
    s.hdr = set_.hdr != 0;
    return s;
  }
  std::string warnings() const { return file_ ? ne_scene_file_warnings(file_.get()) : ""; }

 private:
  std::string dir_;
  int flags_;
  std::shared_ptr<ne_scene_file> file_;
  ne_scene_desc desc_{};
  ne_camera_desc cam_{};
  ne_scene_settings set_{};
};

// One device context (one HIP stream; calls on it are serialised).
class Context {
 public:
  explicit Context(int device = 0) {
    ne_ctx* c = nullptr;
    detail::check(nullptr, ne_ctx_create(device, &c));
    h_.reset(c, detail::CtxDel());
  }
  ne_ctx* get() const { return h_.get(); }
  void synchronize() const { detail::check(get(), ne_synchronize(get())); }

 private:
  std::shared_ptr<ne_ctx> h_;
};

// Uploaded scene (SceneReader's Scene after flattening; see ne_scene_desc).
class Scene {
 public:
  Scene(const Context& ctx, const ne_scene_desc& desc) : ctx_(ctx) {
    ne_scene* s = nullptr;
    detail::check(ctx.get(), ne_scene_upload(ctx.get(), &desc, &s));
    h_.reset(s, detail::SceneDel());
  }
  ne_scene* get() const { return h_.get(); }
  const Context& context() const { return ctx_; }

 private:
  Context ctx_;
  std::shared_ptr<ne_scene> h_;
};

// Integrator::Li on the GPU: one path per ray, FP32, corrected semantics.
// Path i of a call is keyed by (seed, pixel = i, sample = sample_index).
class GpuPathIntegrator {
 public:
  GpuPathIntegrator(const Scene& scene, int bounces, uint64_t seed = 0)
      : scene_(scene), bounces_(bounces), seed_(seed) {}
  // Integrator::Li(Ray incoming, Scene* scene) -- batch of one (slow debug path)
  Vec3 Li(const Ray& incoming) { return Li(std::vector<Ray>{incoming})[0]; }
  std::vector<Vec3> Li(const std::vector<Ray>& rays) {
    std::vector<float> in(rays.size() * 6), out(rays.size() * 3);
    for (size_t i = 0; i < rays.size(); ++i) {
      const Ray& r = rays[i];
      float* p = &in[6 * i];
      p[0] = r.o.x; p[1] = r.o.y; p[2] = r.o.z; p[3] = r.d.x; p[4] = r.d.y; p[5] = r.d.z;
    }
    const ne_ctx* c = scene_.context().get();
    detail::check(c, ne_li(scene_.context().get(), scene_.get(), int32_t(rays.size()), in.data(), bounces_, seed_,
                           sample_++, out.data()));
    std::vector<Vec3> res(rays.size());
    for (size_t i = 0; i < rays.size(); ++i) res[i] = Vec3{out[3 * i], out[3 * i + 1], out[3 * i + 2]};
    return res;
  }
  // Integrator::clone: the copy shares the scene and continues the sample index
  GpuPathIntegrator* clone() const { return new GpuPathIntegrator(*this); }

 private:
  Scene scene_;
  int bounces_;
  uint64_t seed_;
  uint32_t sample_ = 0;
};

// OfflineEngine on the GPU: same tiles, same pixel layout (row-major, x
// fastest, 3 floats per pixel, post-processed).  renderTile may be called
// from many threads; the context serialises them.
class OfflineEngine {
 public:
  static constexpr int kTilesX = 40, kTilesY = 10;  // OfflineEngine.cpp:19

  OfflineEngine(const CameraSpec& camera, const SceneSettings& settings, const Scene& scene, uint64_t seed = 0)
      : camera(camera), settings(settings), scene_(scene), seed_(seed),
        pixels(size_t(settings.width) * size_t(settings.height) * 3, 0.0f) {
    ne_film* f = nullptr;
    detail::check(ctx(), ne_film_create(scene.context().get(), settings.width, settings.height, &f));
    film_.reset(f, detail::FilmDel());
  }
  // OfflineEngine::renderTile (src/core/OfflineEngine.cpp:54-76)
  void renderTile(const CameraSpec& cam, int index, std::atomic<bool>& finished) {
    ne_render_params p = params();
    const ne_camera_desc cd = cam.desc();
    detail::check(ctx(), ne_render_tile(scene_.context().get(), scene_.get(), &cd, &p, index, film_.get(),
                                        pixels.data()));
    finished = true;
  }
  // OfflineEngine::coreLoop's rendering part (src/core/OfflineEngine.cpp:78-145):
  // every tile, in index order
  void coreLoop() {
    std::atomic<bool> done{false};
    for (int i = 0; i < kTilesX * kTilesY; ++i) renderTile(camera, i, done);
  }
  // Whole frame in one launch sequence (the GPU-native path): every pixel,
  // including the remainder columns/rows the 40 x 10 tiling drops.
  void renderFrame() {
    ne_render_params p = params();
    const ne_camera_desc cd = camera.desc();
    detail::check(ctx(), ne_film_clear(film_.get()));
    detail::check(ctx(), ne_render(scene_.context().get(), scene_.get(), &cd, &p, film_.get()));
    detail::check(ctx(), ne_film_tonemap(film_.get(), settings.spp, pixels.data()));
  }
  // linear radiance sums (FP64, 3 per pixel), before post-processing
  std::vector<double> linearSums() const {
    std::vector<double> out(pixels.size());
    detail::check(ctx(), ne_film_read(film_.get(), out.data()));
    return out;
  }

  CameraSpec camera;
  SceneSettings settings;

 private:
  const ne_ctx* ctx() const { return scene_.context().get(); }
  ne_render_params params() const {
    ne_render_params p{};
    p.width = settings.width;
    p.height = settings.height;
    p.x0 = 0; p.y0 = 0; p.x1 = settings.width; p.y1 = settings.height;
    p.spp = settings.spp;
    p.max_bounces = settings.bounces;
    p.seed = seed_;
    p.precision = NE_FP32;
    return p;
  }
  Scene scene_;
  uint64_t seed_;
  std::shared_ptr<ne_film> film_;

 public:
  std::vector<float> pixels;  // glm::vec3 pixels[W*H] of the reference, flattened
};

}  // namespace narval_gpu

#endif  // NE_GPU_HPP
