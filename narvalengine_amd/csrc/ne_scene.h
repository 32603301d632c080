// ne_scene.h -- device-resident scene layout (POD, shared by host upload
// code and kernels).
//
// HBM layout (all arrays 16-byte aligned):
//   nodes   4-wide BVH, 128 B per node = 8 x float4, the children's boxes
//           as structure of arrays:
//             f0 = min.x of children 0..3   f1 = min.y   f2 = min.z
//             f3 = max.x                    f4 = max.y   f5 = max.z
//             f6 = int child[4]             f7 = int count[4]
//           a child whose count word is 0 is an inner node index; otherwise
//           it is a leaf: count = word & 0x0FFFFFFF primitives of type
//           word >> 28 (0 triangle, 1 sphere, 2 rectangle) at slots
//           [index, index + count) of that type's array.  index < 0 marks an
//           empty slot.
//   tris    48 B per triangle in leaf order: (v0, material), (v1-v0, original
//           index), (v2-v0, 0).  Original index = position in the input.
//   sph     (centre, radius) in leaf order + material and original-index
//           arrays; rect 4 x float4 in leaf order:
//           (p0, material), (eu, area), (ev, original index), (n, 0).
//   mats    12 floats per material (see ne_gpu.h).
//   grid    R32F x-fastest density; maj: mres^3 floats, sigma_t majorant per
//           cell (grid value max over the cell's trilinear support x sigma_bar).
#pragma once
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
constexpr uint32_t kItemMaxLeaf = 8;  // count - 1 in 3 bits

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
  float inv[12];          // WCS -> OCS, rows of a 3x4 matrix
  const float* grid;
  const float* maj;
  float sigma_bar;        // avg(sigma_t * densityMultiplier)
  float sig_s[3];
  float sig_t[3];
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
