// bvh_build.cpp -- host binned-SAH BVH builder producing the device layout
// described in ne_scene.h (BVH2, both child boxes stored in the parent,
// leaves referenced from the parent, preorder so a left inner child sits
// right after its parent).
//
// Replaces the reference's midpoint-split builder (src/primitives/BVH.cpp:6-106);
// traversal results (closest primitive, t) are the contract, not the tree
// shape.  Leaves hold one primitive type, so a leaf child points straight at
// that type's storage (no per-primitive indirection in traversal).
#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace ne {

namespace {

struct Box {
  float mn[3], mx[3];
  void reset() {
    for (int a = 0; a < 3; ++a) { mn[a] = std::numeric_limits<float>::infinity(); mx[a] = -mn[a]; }
  }
  void grow(const Box& b) {
    for (int a = 0; a < 3; ++a) { mn[a] = std::min(mn[a], b.mn[a]); mx[a] = std::max(mx[a], b.mx[a]); }
  }
  void grow_pt(const float* p) {
    for (int a = 0; a < 3; ++a) { mn[a] = std::min(mn[a], p[a]); mx[a] = std::max(mx[a], p[a]); }
  }
  float area() const {
    float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    if (!(dx >= 0) || !(dy >= 0) || !(dz >= 0)) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
  }
};

struct BNode {
  Box box;
  int32_t left = -1, right = -1;  // build-node indices (inner) ; -1 for leaves
  uint32_t start = 0, count = 0;  // leaf range in the ref array
  int depth = 0;
};

constexpr int kBins = 16;
constexpr int kMaxLeaf = 4;
constexpr int kMaxDepth = 60;
constexpr float kCostTrav = 1.0f, kCostIsect = 1.0f;

}  // namespace

void build_bvh(const std::vector<PrimInput>& prims, BvhOutput& out) {
  out.nodes.clear();
  out.refs.clear();
  out.max_depth = 0;
  out.max_leaf = 0;
  const size_t n = prims.size();
  if (n == 0) return;
  std::vector<uint32_t> idx(n);
  for (size_t i = 0; i < n; ++i) idx[i] = uint32_t(i);
  std::vector<Box> pbox(n);
  std::vector<float> cen(3 * n);
  for (size_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      pbox[i].mn[a] = prims[i].bmin[a];
      pbox[i].mx[a] = prims[i].bmax[a];
      cen[3 * i + a] = 0.5f * (prims[i].bmin[a] + prims[i].bmax[a]);
    }
  }
  std::vector<BNode> bn;
  bn.reserve(2 * n / kMaxLeaf + 16);
  struct Task { int32_t node; uint32_t start, end; int depth; };
  std::vector<Task> tasks;
  bn.emplace_back();
  tasks.push_back({0, 0, uint32_t(n), 0});
  std::vector<uint32_t> tmp;
  while (!tasks.empty()) {
    Task tk = tasks.back();
    tasks.pop_back();
    BNode& nd = bn[size_t(tk.node)];
    nd.depth = tk.depth;
    Box b, cb;
    b.reset();
    cb.reset();
    for (uint32_t i = tk.start; i < tk.end; ++i) {
      b.grow(pbox[idx[i]]);
      cb.grow_pt(&cen[3 * size_t(idx[i])]);
    }
    nd.box = b;
    const uint32_t cnt = tk.end - tk.start;
    auto make_leaf = [&]() {
      BNode& m = bn[size_t(tk.node)];
      m.start = tk.start;
      m.count = cnt;
      m.left = m.right = -1;
    };
    // a leaf over a range that mixes primitive types is split by type instead
    auto leaf_or_type_split = [&]() {
      const uint32_t t0 = prims[idx[tk.start]].type;
      bool mixed = false;
      for (uint32_t i = tk.start + 1; i < tk.end && !mixed; ++i) mixed = prims[idx[i]].type != t0;
      if (!mixed) { make_leaf(); return; }
      auto it = std::stable_partition(idx.begin() + tk.start, idx.begin() + tk.end,
                                      [&](uint32_t p) { return prims[p].type == t0; });
      const uint32_t mid = uint32_t(it - idx.begin());
      int32_t l = int32_t(bn.size());
      bn.emplace_back();
      int32_t r = int32_t(bn.size());
      bn.emplace_back();
      bn[size_t(tk.node)].left = l;
      bn[size_t(tk.node)].right = r;
      tasks.push_back({r, mid, tk.end, tk.depth + 1});
      tasks.push_back({l, tk.start, mid, tk.depth + 1});
    };
    if (cnt <= 1 || tk.depth >= kMaxDepth) { leaf_or_type_split(); continue; }
    // binned SAH over the three axes
    float best_cost = std::numeric_limits<float>::infinity();
    int best_axis = -1, best_split = -1;
    const float parent_area = b.area();
    for (int a = 0; a < 3; ++a) {
      float lo = cb.mn[a], ext = cb.mx[a] - cb.mn[a];
      if (!(ext > 0)) continue;
      Box bins[kBins];
      uint32_t counts[kBins] = {0};
      for (int k = 0; k < kBins; ++k) bins[k].reset();
      const float scale = float(kBins) / ext;
      for (uint32_t i = tk.start; i < tk.end; ++i) {
        uint32_t p = idx[i];
        int k = int((cen[3 * size_t(p) + a] - lo) * scale);
        k = std::min(std::max(k, 0), kBins - 1);
        counts[k]++;
        bins[k].grow(pbox[p]);
      }
      float right_area[kBins];
      uint32_t right_count[kBins];
      Box acc;
      acc.reset();
      uint32_t c = 0;
      for (int k = kBins - 1; k > 0; --k) {
        acc.grow(bins[k]);
        c += counts[k];
        right_area[k] = acc.area();
        right_count[k] = c;
      }
      acc.reset();
      c = 0;
      for (int k = 0; k < kBins - 1; ++k) {
        acc.grow(bins[k]);
        c += counts[k];
        if (c == 0 || right_count[k + 1] == 0) continue;
        float cost = kCostTrav + kCostIsect * (acc.area() * float(c) + right_area[k + 1] * float(right_count[k + 1])) /
                                     std::max(parent_area, 1e-30f);
        if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = k; }
      }
    }
    const float leaf_cost = kCostIsect * float(cnt);
    if (cnt <= uint32_t(kMaxLeaf) && (best_axis < 0 || leaf_cost <= best_cost)) { leaf_or_type_split(); continue; }
    uint32_t mid;
    if (best_axis < 0) {
      // degenerate centroids: split the range in half
      mid = tk.start + cnt / 2;
    } else {
      const float lo = cb.mn[best_axis], scale = float(kBins) / (cb.mx[best_axis] - cb.mn[best_axis]);
      auto it = std::partition(idx.begin() + tk.start, idx.begin() + tk.end, [&](uint32_t p) {
        int k = int((cen[3 * size_t(p) + best_axis] - lo) * scale);
        k = std::min(std::max(k, 0), kBins - 1);
        return k <= best_split;
      });
      mid = uint32_t(it - idx.begin());
      if (mid == tk.start || mid == tk.end) mid = tk.start + cnt / 2;
    }
    int32_t l = int32_t(bn.size());
    bn.emplace_back();
    int32_t r = int32_t(bn.size());
    bn.emplace_back();
    bn[size_t(tk.node)].left = l;
    bn[size_t(tk.node)].right = r;
    tasks.push_back({r, mid, tk.end, tk.depth + 1});
    tasks.push_back({l, tk.start, mid, tk.depth + 1});
  }

  // ---- emit: refs in leaf order, inner nodes in preorder ------------------
  out.order.clear();
  out.order.reserve(n);
  std::vector<int32_t> out_index(bn.size(), -1);
  // first pass: preorder numbering of inner nodes
  std::vector<int32_t> st;
  std::vector<int32_t> pre;
  st.push_back(0);
  while (!st.empty()) {
    int32_t k = st.back();
    st.pop_back();
    const BNode& nd = bn[size_t(k)];
    if (nd.left < 0) continue;
    out_index[size_t(k)] = int32_t(pre.size());
    pre.push_back(k);
    st.push_back(nd.right);
    st.push_back(nd.left);
  }
  const bool root_leaf = bn[0].left < 0;
  size_t n_out = root_leaf ? 1 : pre.size();
  out.nodes.assign(n_out * 16, 0.0f);
  auto put_box = [&](float* f, int side, const Box& b) {
    // side 0: f0 = (lmin, lmax.x) f1.xy = lmax.yz ; side 1: f1.zw = rmin.xy, f2 = (rmin.z, rmax)
    if (side == 0) {
      f[0] = b.mn[0]; f[1] = b.mn[1]; f[2] = b.mn[2]; f[3] = b.mx[0]; f[4] = b.mx[1]; f[5] = b.mx[2];
    } else {
      f[6] = b.mn[0]; f[7] = b.mn[1]; f[8] = b.mn[2]; f[9] = b.mx[0]; f[10] = b.mx[1]; f[11] = b.mx[2];
    }
  };
  auto put_int = [&](float* f, int slot, int32_t v) { std::memcpy(&f[12 + slot], &v, 4); };
  auto emit_leaf = [&](const BNode& leaf, float* f, int side) {
    int32_t first = int32_t(out.refs.size());
    for (uint32_t i = leaf.start; i < leaf.start + leaf.count; ++i) {
      uint32_t p = idx[i];
      out.order.push_back(p);
      out.refs.push_back(p);  // patched to (type, storage) by the caller
    }
    put_box(f, side, leaf.box);
    put_int(f, side, first);
    put_int(f, 2 + side, int32_t(leaf.count | (prims[idx[leaf.start]].type << kLeafTypeShift)));
    out.max_leaf = std::max(out.max_leaf, leaf.count);
  };
  if (root_leaf) {
    float* f = &out.nodes[0];
    emit_leaf(bn[0], f, 0);
    Box e;
    e.reset();
    put_box(f, 1, e);
    put_int(f, 1, -1);
    put_int(f, 3, 0);
    out.max_depth = 1;
    return;
  }
  for (size_t j = 0; j < pre.size(); ++j) {
    const BNode& nd = bn[size_t(pre[j])];
    float* f = &out.nodes[j * 16];
    out.max_depth = std::max(out.max_depth, uint32_t(nd.depth + 1));
    const int32_t ch[2] = {nd.left, nd.right};
    for (int side = 0; side < 2; ++side) {
      const BNode& c = bn[size_t(ch[side])];
      if (c.left < 0) {
        emit_leaf(c, f, side);
      } else {
        put_box(f, side, c.box);
        put_int(f, side, out_index[size_t(ch[side])]);
        put_int(f, 2 + side, 0);
      }
    }
  }
}

}  // namespace ne
