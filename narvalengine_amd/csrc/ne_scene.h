// ne_scene.h -- device-resident scene layout (POD, shared by host upload
// code and kernels).
//
// HBM layout (all arrays 16-byte aligned):
//   nodes   4-wide BVH, 64 B per node = 4 x float4, child boxes quantized
//           to 8 bits per plane against the node's box:
//             f0 = (origin.xyz, exponents: byte a = e_a + 127)
//             f1 = (qlo.x, qlo.y, qlo.z, qhi.x)  f2 = (qhi.y, qhi.z, meta, 0)
//             f3 = int child[4]
//           each q word holds one byte per child; child c's box on axis a is
//           [origin_a + qlo_a[c] * 2^e_a, origin_a + qhi_a[c] * 2^e_a] in FP32
//           (the builder makes that decode contain the true box).  meta byte
//           c is 0 for an inner child (child[c] = node index) or
//           0x80 | type << 4 | count for a leaf of count primitives of type
//           (0 triangle, 1 sphere, 2 rectangle) at slots [child[c],
//           child[c] + count) of that type's array.  child < 0: empty slot.
//   tris    48 B per triangle in leaf order: (v0, material), (v1-v0, original
//           index), (v2-v0, 0).  Original index = position in the input.
//   sph     (centre, radius) in leaf order + material and original-index
//           arrays; rect 4 x float4 in leaf order:
//           (p0, material), (eu, area), (ev, original index), (n, 0).
//   mats    12 floats per material (see ne_gpu.h).
//   grid    R32F density in 4 x 4 x 2 voxel bricks (grid_offset below);
//           maj: mres^3 floats, sigma_t majorant per cell (grid value max over
//           the cell's trilinear support x sigma_bar).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ne {

enum { PRIM_TRI = 0, PRIM_SPHERE = 1, PRIM_RECT = 2 };
constexpr int kLightPoint = 3;  // light-code type of a point light (not a BVH primitive)

// BVH traversal stack: kShortStack entries per lane in LDS ([depth][lane]
// layout, no bank conflicts), deeper entries spill to a private array.  Upload
// rejects a tree whose deepest possible stack (children pushed along any root
// path) exceeds the total.
constexpr int kShortStack = 16;
constexpr int kSpillStack = 48;

// Traversal work item: an inner node index, or a leaf packed as
// kItemLeaf | type << 29 | (count - 1) << 26 | first slot.
constexpr uint32_t kItemLeaf = 0x80000000u;
constexpr uint32_t kItemFirstMask = (1u << 26) - 1;  // slots per primitive type < 2^26
constexpr uint32_t kItemCountShift = 26;
constexpr uint32_t kItemNone = 0xFFFFFFFFu;  // no item (leaf type 3 does not exist)
constexpr uint32_t kItemMaxLeaf = 8;  // count - 1 in 3 bits

// Density grid layout: bricks of 4 x 4 x 2 voxels (32 floats: one 128-byte
// cache line), stored x-fastest, then y, then z.  A trilinear lookup's 2 x 2 x 2
// corners then span 2.3 lines on average instead of 4 in a row-major grid.
// The grid is padded with zeros to grid_pad_{x,y,z}(res) >= res + 2 voxels per
// axis: corners at an index >= res read 0 without a bounds check.
constexpr uint32_t kBrickFloats = 32;
__host__ __device__ inline uint32_t grid_pad_x(int32_t r) { return (uint32_t(r) + 2u + 3u) & ~3u; }
__host__ __device__ inline uint32_t grid_pad_y(int32_t r) { return (uint32_t(r) + 2u + 3u) & ~3u; }
__host__ __device__ inline uint32_t grid_pad_z(int32_t r) { return (uint32_t(r) + 2u + 1u) & ~1u; }
__host__ __device__ inline uint64_t grid_offset(uint32_t x, uint32_t y, uint32_t z, uint32_t step_y, uint32_t step_z) {
  return uint64_t(z >> 1) * step_z + uint64_t(y >> 2) * step_y + uint64_t(x >> 2) * kBrickFloats +
         (((z & 1u) << 4) | ((y & 3u) << 2) | (x & 3u));
}

struct DevScene {
  const float4* nodes;
  const float4* tris;
  const float4* sph;
  const int32_t* sph_mat;
  const int32_t* sph_orig;
  const float4* rects;
  const float* mats;
  const uint32_t* emitters;  // emissive primitives: type << 30 | storage slot
  const float4* points;      // point lights: position, material (int bits)
  uint32_t n_nodes;
  int32_t n_emitters;
  // volume
  int32_t has_volume;
  int32_t homog;          // homogeneous medium: density 1 everywhere in the cube
  int32_t res[3];
  int32_t mres;
  int32_t occ_lo[3], occ_hi[3];  // majorant cells with a positive majorant lie in [occ_lo, occ_hi]
                                 // (occ_lo > occ_hi: none)
  float inv[12];          // WCS -> OCS, rows of a 3x4 matrix
  const float* grid;      // bricked (grid_offset), zero-padded to at least res + 2 voxels per axis
  uint32_t grid_step_y;   // floats from one brick row (4 voxels in y) to the next
  uint32_t grid_step_z;   // floats from one brick slab (2 voxels in z) to the next
  int32_t grid_small;     // the bricked grid spans under 2^32 bytes and grid_step_z < 2^24: 32-bit byte offsets address it
  const float* maj;
  const float2* maj_pair;  // per cell {maj value, 1 / maj (correctly rounded; 0 unless maj is positive and finite)}
  float sigma_bar;        // avg(sigma_t * densityMultiplier)
  float sig_s[3];
  float sig_t[3];
  float albedo[3];  // sig_s / sig_t, divided once at upload (IEEE division, as on the device)
  float g;
  // lights
  int32_t has_sun;
  float sun_dir[3];
  float sun_E[3];
  int32_t env_type;
  float env[9];
  // environment map (env_type 3): RGB texels, per-texel sampling function,
  // per-row conditional CDFs (W + 1 each), marginal CDF (H + 1)
  const float* env_tex;
  const float* env_func;
  const float* env_cdf;
  const float* env_marg;
  int32_t env_w, env_h;
  float env_inv_int;  // 1 / (mean of env_func)
  int32_t n_lights;
};

struct DevCamera {
  float pos[3], lower_left[3], horizontal[3], vertical[3];
};

struct DevCounters {
  unsigned long long paths, ext_rays, shadow_rays, density_lookups, majorant_cells, bvh_nodes, prim_tests;
  unsigned long long sh_lookups, sh_cells, sh_nodes, sh_prims;  // shadow-ray share (wavefront)
  unsigned long long guide_queries;                             // learned-proposal evaluations
};

}  // namespace ne
