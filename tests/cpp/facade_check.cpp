// facade_check.cpp -- drives the C++ facade (include/ne_gpu.hpp) the way the
// reference drives its renderer: OfflineEngine tiles, a whole frame, and
// Integrator::Li.  Exits non-zero on any inconsistency:
//   * every tile pixel equals the whole-frame pixel (same counter-RNG paths)
//   * pixels outside the 40 x 10 tile grid stay untouched by the tile calls
//   * Li radiance is finite and non-negative; a ray that misses everything
//     returns the constant environment radiance exactly
//   * SceneReader loads the reference's surfaceAndLight test scene from the
//     resources directory given as argv[1] (host only, before any device
//     call), and on the device its frame renders to finite pixels
// Exit codes: 0 pass, 1 mismatch, 3 no device / device error, 4 reader failure.
#include <cmath>
#include <cstdio>
#include <vector>

#include "ne_gpu.hpp"

int main(int argc, char** argv) {
  using namespace narval_gpu;
  SceneReader reader(argc > 1 ? argv[1] : "");
  if (argc > 1) {
    try {
      reader.loadScene("scenes/testing/surfaceAndLight.json", false);
      const ne_scene_desc& sd = reader.getScene();
      const SceneSettings st = reader.getSettings();
      const CameraSpec mc = reader.getMainCamera();
      if (sd.n_rect != 2 || sd.n_mat != 2 || st.width != 600 || st.height != 300 || st.bounces != 6 ||
          mc.aspect != 2.0f || mc.up.y != 1.0f || sd.rect[12 + 10] != 1.0f) {
        std::fprintf(stderr, "scene reader: unexpected content\n");
        return 4;
      }
    } catch (const Error& e) {
      std::fprintf(stderr, "scene reader error %d: %s\n", int(e.status), e.what());
      return 4;
    }
  }
  try {
    Context ctx(0);
    const int res = 32;
    std::vector<float> grid(size_t(res) * res * res);
    if (ne_gen_fbm_grid(res, 0, 4.0f, grid.data()) != NE_OK) return 2;
    ne_scene_desc d{};
    d.has_volume = 1;
    const float id[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    for (int k = 0; k < 16; ++k) { d.vol_xf[k] = id[k]; d.vol_inv[k] = id[k]; }
    d.vol_res[0] = d.vol_res[1] = d.vol_res[2] = res;
    d.grid = grid.data();
    d.vol_mult = 20.0f;
    for (int c = 0; c < 3; ++c) { d.vol_sigma_s[c] = 0.99f; d.vol_sigma_a[c] = 0.01f; }
    d.vol_g = 0.85f;
    d.maj_res = 4;
    d.has_sun = 1;
    const float sd[3] = {-0.4082483f, -0.8164966f, 0.2857738f};
    for (int c = 0; c < 3; ++c) { d.sun_dir[c] = sd[c]; d.sun_E[c] = 1.0f; }
    d.env_type = NE_ENV_CONSTANT;
    d.env_c0[0] = 0.2f; d.env_c0[1] = 0.25f; d.env_c0[2] = 0.35f;
    Scene scene(ctx, d);

    CameraSpec cam;
    cam.position = {0.0f, 0.25f, -1.8f};
    cam.look_at = {0.0f, 0.0f, 0.0f};
    cam.vfov = 45.0f;
    SceneSettings set;
    set.width = 84;   // 84 / 40 = 2: columns 80..83 are not covered by tiles
    set.height = 43;  // 43 / 10 = 4: rows 40..42 are not covered
    set.spp = 8;
    set.bounces = 16;

    OfflineEngine tiles(cam, set, scene, 7);
    for (auto& p : tiles.pixels) p = -1.0f;
    tiles.coreLoop();
    OfflineEngine frame(cam, set, scene, 7);
    frame.renderFrame();
    int bad = 0, covered = 0;
    const int tw = set.width / 40, th = set.height / 10;
    for (int y = 0; y < set.height; ++y)
      for (int x = 0; x < set.width; ++x)
        for (int c = 0; c < 3; ++c) {
          const size_t i = (size_t(y) * set.width + x) * 3 + c;
          if (x < 40 * tw && y < 10 * th) {
            ++covered;
            if (tiles.pixels[i] != frame.pixels[i]) ++bad;
          } else if (tiles.pixels[i] != -1.0f) {
            ++bad;
          }
        }
    // a tile rendered twice is recomputed, not accumulated
    std::vector<float> before = tiles.pixels;
    std::atomic<bool> done{false};
    tiles.renderTile(cam, 123, done);
    if (!done || tiles.pixels != before) ++bad;

    GpuPathIntegrator integ(scene, 16, 3);
    std::vector<Ray> rays;
    rays.push_back(Ray{{0.0f, 0.25f, -1.8f}, {0.0f, -0.1373606f, 0.9905211f}});
    rays.push_back(Ray{{0.0f, 5.0f, 0.0f}, {0.0f, 1.0f, 0.0f}});  // misses the volume
    std::unique_ptr<GpuPathIntegrator> twin(integ.clone());
    const std::vector<Vec3> L = integ.Li(rays);
    const Vec3 L1 = twin->Li(rays[0]);
    for (const Vec3& v : L)
      if (!(std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z)) || v.x < 0 || v.y < 0 || v.z < 0) ++bad;
    if (L[1].x != 0.2f || L[1].y != 0.25f || L[1].z != 0.35f) ++bad;
    if (!std::isfinite(L1.x)) ++bad;
    if (argc > 1) {  // the loaded reference scene renders on the device
      Scene loaded(ctx, reader.getScene());
      SceneSettings ls = reader.getSettings();
      ls.width = 120;
      ls.height = 60;
      ls.spp = 4;
      OfflineEngine eng(reader.getMainCamera(), ls, loaded, 1);
      eng.renderFrame();
      double sum = 0;
      for (float v : eng.pixels) {
        if (!std::isfinite(v) || v < 0 || v > 1) ++bad;
        sum += v;
      }
      if (!(sum > 0)) ++bad;
    }
    std::printf("{\"covered\": %d, \"mismatches\": %d, \"L0\": [%g, %g, %g]}\n", covered, bad, L[0].x, L[0].y,
                L[0].z);
    return bad == 0 ? 0 : 1;
  } catch (const Error& e) {
    std::fprintf(stderr, "narval_gpu error %d: %s\n", int(e.status), e.what());
    return 3;
  }
}
