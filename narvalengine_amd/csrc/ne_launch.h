// ne_launch.h -- host-callable launchers of the HIP kernels (C++ linkage,
// internal to the library; the public surface is include/ne_gpu.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ne_scene.h"

namespace ne {

struct RenderArgs {
  int32_t W, H, x0, y0, x1, y1;
  int32_t spp, max_bounces;
  uint32_t sample_offset, job;
  uint64_t seed;
  int32_t S;        // samples per thread
  int32_t nchunks;  // ceil(spp / S)
  int32_t ntx, nty; // 8x8 pixel tiles covering the rect
  double* out;      // film (nchunks == 1) or partial sums [nchunks][rect pixels][3]
  DevCounters* counters;
};

hipError_t launch_render(const DevScene& s, const DevCamera& cam, const RenderArgs& a, int precision, bool count,
                         int block, hipStream_t st);
hipError_t launch_reduce_partials(const RenderArgs& a, double* film, hipStream_t st);
hipError_t launch_tonemap(const double* film, int npix, float inv_spp, float* pixels, hipStream_t st);

hipError_t launch_test_primitive(int kind, int n, const float* rays, const float* prims, int32_t* hit, float* t0,
                                 float* t1, hipStream_t st);
hipError_t launch_test_closest(const DevScene& s, int n, const float* rays, float tmin, float tmax, float* t,
                               int32_t* type, int32_t* index, hipStream_t st);
hipError_t launch_test_transmittance(const DevScene& s, int n, const float* rays, const float* tmax, uint64_t seed,
                                     int precision, double* out, hipStream_t st);
hipError_t launch_test_delta(const DevScene& s, int n, const float* rays, uint64_t seed, int precision,
                             int32_t* scattered, double* t, hipStream_t st);
hipError_t launch_test_density(const DevScene& s, int n, const float* gcs, float* out, hipStream_t st);
hipError_t launch_test_bsdf(int kind, int n, const float* in0, const float* in1, const float* in2, const float* in3,
                            float* out, hipStream_t st);
hipError_t launch_test_rng(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t job, int n, float* out,
                           hipStream_t st);
hipError_t launch_li(const DevScene& s, int n, const float* rays, int max_bounces, uint64_t seed, uint32_t sample,
                     float* out, hipStream_t st);
hipError_t launch_majorants(const DevScene& s, float* maj, hipStream_t st);

}  // namespace ne
