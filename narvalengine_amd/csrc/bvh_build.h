// bvh_build.h -- host SAH builder interface.
#pragma once
#include <stdint.h>

#include <vector>

namespace ne {

struct PrimInput {
  float bmin[3], bmax[3];
  uint32_t type;  // leaves never mix types
};

constexpr int kNodeFloats = 16;   // 4-wide quantized node: 4 x float4 (64 B)
constexpr uint32_t kMetaLeaf = 0x80;  // child meta byte: 0x80 | type << 4 | count for a leaf, 0 otherwise

struct BvhOutput {
  std::vector<float> nodes;     // kNodeFloats per node, ne_scene.h layout
  std::vector<uint32_t> refs;   // leaf slots -> input primitive index (leaves hold refs[first, first+count))
  std::vector<uint32_t> order;  // input primitive indices in leaf order
  uint32_t max_depth = 0;       // wide-node levels
  uint32_t max_stack = 0;       // deepest traversal stack (children pushed along any root path)
  uint32_t max_leaf = 0;
  bool quant_failed = false;    // some child box has no conservative 8-bit encoding (non-finite bounds)
};

void build_bvh(const std::vector<PrimInput>& prims, BvhOutput& out);

}  // namespace ne
