// bvh_build.cpp -- host binned-SAH BVH builder producing the device layout
// described in ne_scene.h: a binary SAH tree collapsed to 4-wide nodes that
// store their children's boxes, leaves referenced from the parent, wide
// nodes in preorder.
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
  out.max_stack = 0;
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

  // ---- collapse to a 4-wide tree and emit -----------------------------------
  // A wide node starts from its binary node's two children and opens the
  // inner child of largest surface area until it has four children or only
  // leaves.  Wide nodes are numbered in preorder; leaves are referenced from
  // their parent's child slots, their primitive refs stored contiguously.
  struct Wide {
    int32_t bnode;
    int32_t ch[4];
    int n;
  };
  auto collapse = [&](int32_t k, Wide& w) {
    w.bnode = k;
    w.n = 0;
    w.ch[w.n++] = bn[size_t(k)].left;
    w.ch[w.n++] = bn[size_t(k)].right;
    while (w.n < 4) {
      int best = -1;
      float best_area = -1.0f;
      for (int i = 0; i < w.n; ++i) {
        const BNode& c = bn[size_t(w.ch[i])];
        if (c.left >= 0 && c.box.area() > best_area) { best_area = c.box.area(); best = i; }
      }
      if (best < 0) break;
      const int32_t open = w.ch[best];
      w.ch[best] = bn[size_t(open)].left;
      w.ch[w.n++] = bn[size_t(open)].right;
    }
  };
  out.order.clear();
  out.order.reserve(n);
  std::vector<Wide> wide;
  std::vector<int32_t> out_index(bn.size(), -1);
  const bool root_leaf = bn[0].left < 0;
  if (!root_leaf) {
    struct Item { int32_t k; uint32_t depth, pushes; };
    std::vector<Item> st;
    st.push_back({0, 1, 0});
    while (!st.empty()) {
      const Item it = st.back();
      st.pop_back();
      Wide w;
      collapse(it.k, w);
      out_index[size_t(it.k)] = int32_t(wide.size());
      wide.push_back(w);
      // a visit of this node pushes at most n - 1 children on the traversal stack
      const uint32_t pushes = it.pushes + uint32_t(w.n - 1);
      out.max_depth = std::max(out.max_depth, it.depth);
      out.max_stack = std::max(out.max_stack, pushes);
      for (int c = w.n - 1; c >= 0; --c)
        if (bn[size_t(w.ch[c])].left >= 0) st.push_back({w.ch[c], it.depth + 1, pushes});
    }
  }
  const size_t n_out = root_leaf ? 1 : wide.size();
  out.nodes.assign(n_out * kNodeFloats, 0.0f);
  auto put_box = [&](float* f, int c, const Box& b) {
    f[c] = b.mn[0]; f[4 + c] = b.mn[1]; f[8 + c] = b.mn[2];
    f[12 + c] = b.mx[0]; f[16 + c] = b.mx[1]; f[20 + c] = b.mx[2];
  };
  auto put_int = [&](float* f, int slot, int32_t v) { std::memcpy(&f[slot], &v, 4); };
  auto emit_leaf = [&](const BNode& leaf, float* f, int c) {
    int32_t first = int32_t(out.refs.size());
    for (uint32_t i = leaf.start; i < leaf.start + leaf.count; ++i) {
      uint32_t p = idx[i];
      out.order.push_back(p);
      out.refs.push_back(p);  // patched to (type, storage) by the caller
    }
    put_box(f, c, leaf.box);
    put_int(f, 24 + c, first);
    put_int(f, 28 + c, int32_t(leaf.count | (prims[idx[leaf.start]].type << kLeafTypeShift)));
    out.max_leaf = std::max(out.max_leaf, leaf.count);
  };
  auto empty_slot = [&](float* f, int c) {
    Box e;
    e.reset();
    put_box(f, c, e);
    put_int(f, 24 + c, -1);
    put_int(f, 28 + c, 0);
  };
  if (root_leaf) {
    float* f = &out.nodes[0];
    emit_leaf(bn[0], f, 0);
    for (int c = 1; c < 4; ++c) empty_slot(f, c);
    out.max_depth = 1;
    out.max_stack = 0;
    return;
  }
  for (size_t j = 0; j < wide.size(); ++j) {
    const Wide& w = wide[j];
    float* f = &out.nodes[j * kNodeFloats];
    for (int c = 0; c < 4; ++c) {
      if (c >= w.n) { empty_slot(f, c); continue; }
      const BNode& ch = bn[size_t(w.ch[c])];
      if (ch.left < 0) {
        emit_leaf(ch, f, c);
      } else {
        put_box(f, c, ch.box);
        put_int(f, 24 + c, out_index[size_t(w.ch[c])]);
        put_int(f, 28 + c, 0);
      }
    }
  }
}

}  // namespace ne
