// bvh_build.cpp -- host binned-SAH BVH builder producing the device layout
// described in ne_scene.h: a binary SAH tree collapsed to 4-wide nodes that
// store their children's boxes quantized to 8 bits per plane against the
// node's own box (conservatively: the FP32 decode the traversal performs
// always contains the true box), leaves referenced from the parent, wide
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
  struct Child {
    Box box;
    int32_t ref = -1;   // inner node index, leaf first ref, or -1
    uint32_t meta = 0;  // 0 inner/empty; leaf: 0x80 | type << 4 | count
  };
  auto leaf_child = [&](const BNode& leaf) {
    Child c;
    c.box = leaf.box;
    c.ref = int32_t(out.refs.size());
    for (uint32_t i = leaf.start; i < leaf.start + leaf.count; ++i) {
      uint32_t p = idx[i];
      out.order.push_back(p);
      out.refs.push_back(p);  // patched to (type, storage) by the caller
    }
    c.meta = kMetaLeaf | (prims[idx[leaf.start]].type << 4) | leaf.count;
    out.max_leaf = std::max(out.max_leaf, leaf.count);
    return c;
  };
  auto encode = [&](float* f, const Child* ch, int n) {
    Box parent;
    parent.reset();
    for (int c = 0; c < n; ++c) parent.grow(ch[c].box);
    uint32_t q[6] = {0, 0, 0, 0, 0, 0};  // lo.x lo.y lo.z hi.x hi.y hi.z, one byte per child
    uint32_t exps = 0;
    for (int a = 0; a < 3; ++a) {
      const float origin = parent.mn[a];
      f[a] = origin;
      const double ext = double(parent.mx[a]) - double(origin);
      int e = ext > 0 && std::isfinite(ext) ? int(std::ceil(std::log2(ext / 254.0))) : -126;
      if (!std::isfinite(ext)) e = 127;
      e = std::min(std::max(e, -126), 127);
      // smallest scale 2^e whose decoded planes origin + q * 2^e (FP32, as the
      // traversal computes them) contain every child box
      for (;; ++e) {
        const float sc = std::ldexp(1.0f, e);
        bool ok = true;
        uint32_t lo[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0};
        for (int c = 0; c < n && ok; ++c) {
          const float bl = ch[c].box.mn[a], bh = ch[c].box.mx[a];
          double ql = std::floor((double(bl) - double(origin)) / double(sc));
          double qh = std::ceil((double(bh) - double(origin)) / double(sc));
          int il = int(std::min(255.0, std::max(0.0, ql))), ih = int(std::min(255.0, std::max(0.0, qh)));
          while (il > 0 && origin + float(il) * sc > bl) --il;
          while (ih < 255 && origin + float(ih) * sc < bh) ++ih;
          ok = origin + float(il) * sc <= bl && origin + float(ih) * sc >= bh;
          lo[c] = uint32_t(il);
          hi[c] = uint32_t(ih);
        }
        if (ok || e == 127) {
          // no scale contains the children: only possible with non-finite
          // bounds, which upload rejects; never emit a box that is too small
          if (!ok) out.quant_failed = true;
          for (int c = 0; c < n; ++c) {
            q[a] |= lo[c] << (8 * c);
            q[3 + a] |= hi[c] << (8 * c);
          }
          break;
        }
      }
      exps |= uint32_t(e + 127) << (8 * a);
    }
    uint32_t meta = 0;
    int32_t refs[4] = {-1, -1, -1, -1};
    for (int c = 0; c < n; ++c) {
      meta |= ch[c].meta << (8 * c);
      refs[c] = ch[c].ref;
    }
    std::memcpy(&f[3], &exps, 4);
    std::memcpy(&f[4], &q[0], 16);  // lo.x lo.y lo.z hi.x
    std::memcpy(&f[8], &q[4], 8);   // hi.y hi.z
    std::memcpy(&f[10], &meta, 4);
    std::memcpy(&f[12], refs, 16);
  };
  if (root_leaf) {
    const Child c = leaf_child(bn[0]);
    encode(&out.nodes[0], &c, 1);
    out.max_depth = 1;
    out.max_stack = 0;
    return;
  }
  for (size_t j = 0; j < wide.size(); ++j) {
    const Wide& w = wide[j];
    Child ch[4];
    for (int c = 0; c < w.n; ++c) {
      const BNode& b = bn[size_t(w.ch[c])];
      if (b.left < 0) {
        ch[c] = leaf_child(b);
      } else {
        ch[c].box = b.box;
        ch[c].ref = out_index[size_t(w.ch[c])];
      }
    }
    encode(&out.nodes[j * kNodeFloats], ch, w.n);
  }
}

}  // namespace ne
