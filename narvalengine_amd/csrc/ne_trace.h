// ne_trace.h -- device functions of the path tracer: primitive tests, BVH
// traversal with an LDS short stack, majorant-grid delta/ratio tracking,
// phase/BSDF sampling, and the per-path integrator.
//
// Everything is templated on R (float for the product path, double for the
// converged-oracle mode) and on COUNT (event counters compiled in or out).
#pragma once
#include "ne_math.h"
#include "ne_scene.h"
#include "ne_vmf.h"

namespace ne {

// Event counters held in registers by one thread.
template <bool COUNT>
struct Count {
  NE_DEV void ext() {}
  NE_DEV void shadow() {}
  NE_DEV void lookup() {}
  NE_DEV void cell() {}
  NE_DEV void node() {}
  NE_DEV void prim() {}
};
template <>
struct Count<true> {
  uint32_t n_ext = 0, n_shadow = 0, n_lookup = 0, n_cell = 0, n_node = 0, n_prim = 0;
  NE_DEV void ext() { ++n_ext; }
  NE_DEV void shadow() { ++n_shadow; }
  NE_DEV void lookup() { ++n_lookup; }
  NE_DEV void cell() { ++n_cell; }
  NE_DEV void node() { ++n_node; }
  NE_DEV void prim() { ++n_prim; }
};

// ---------------------------------------------------------------------------
// Reference-semantics primitive tests (FP32), used by the parity entry points.
// ---------------------------------------------------------------------------
// AABB::intersect, src/primitives/AABB.cpp:48-79
NE_DEV bool ref_aabb(Vec3<float> bmin, Vec3<float> bmax, Vec3<float> o, Vec3<float> d, float& tn, float& tf) {
  Vec3<float> s = v3((d.x < 0.0) ? 1.0f : -1.0f, (d.y < 0.0) ? 1.0f : -1.0f, (d.z < 0.0) ? 1.0f : -1.0f);
  Vec3<float> invD = v3s(1.0f) / d;
  Vec3<float> size = bmax - bmin;
  Vec3<float> half = size / 2.0f;
  Vec3<float> center = bmax - size / 2.0f;
  Vec3<float> t1 = (center - o + s * half) * invD;
  Vec3<float> t2 = (center - o - s * half) * invD;
  Vec3<float> mn = v3(xmin(t1.x, t2.x), xmin(t1.y, t2.y), xmin(t1.z, t2.z));
  Vec3<float> mx = v3(xmax(t1.x, t2.x), xmax(t1.y, t2.y), xmax(t1.z, t2.z));
  float a = xmax(xmax(mn.x, mn.y), mn.z);
  float b = xmin(xmin(mx.x, mx.y), mx.z);
  if (a > b) { tn = __builtin_huge_valf(); tf = -__builtin_huge_valf(); return false; }
  tn = a; tf = b;
  return true;
}
// Triangle::intersect, src/primitives/Triangle.cpp:49-81
NE_DEV bool ref_triangle(Vec3<float> v0, Vec3<float> v1, Vec3<float> v2, Vec3<float> o, Vec3<float> d, float& t) {
  Vec3<float> v1v0 = v1 - v0, v2v0 = v2 - v0, rov0 = o - v0;
  Vec3<float> n = cross(v1v0, v2v0);
  Vec3<float> q = cross(rov0, d);
  float denom = float(1.0 / double(dot(d, n)));
  float u = dot(-q, v2v0) * denom;
  float v = dot(q, v1v0) * denom;
  float tt = dot(-n, rov0) * denom;
  t = __builtin_huge_valf();
  if (__builtin_isnan(tt) || tt < 0) return false;
  if (u < 0.0f || u > 1.0f || v < 0.0f || (u + v) > 1.0f) return false;
  t = tt;
  return true;
}
// Sphere::intersect, src/primitives/Sphere.cpp:21-44
NE_DEV bool ref_sphere(Vec3<float> c, float r, Vec3<float> o, Vec3<float> d, float& tn, float& tf) {
  Vec3<float> oc = o - c;
  float a = dot(d, d), b = dot(oc, d), cc = dot(oc, oc) - r * r;
  float disc = b * b - a * cc;
  tn = __builtin_huge_valf(); tf = -__builtin_huge_valf();
  if (disc >= 0) {
    float t1 = (-b - sqrtf(disc)) / a, t2 = (-b + sqrtf(disc)) / a;
    tn = xmin(t1, t2); tf = xmax(t1, t2);
    return true;
  }
  return false;
}
// Rectangle::intersect, src/primitives/Rectangle.cpp:16-76
NE_DEV bool ref_rect(Vec3<float> vA, Vec3<float> vB, Vec3<float> nrm, Vec3<float> o, Vec3<float> d, float& t) {
  t = __builtin_huge_valf();
  float denom = dot(nrm, d);
  if (fabsf(denom) < float(kEPS)) return false;
  float num = dot(nrm, vB - o);
  float tt = num / denom;
  if (tt < 0) return false;
  Vec3<float> p = o + tt * d;
  Vec3<float> a = vA, b = vB;
  if (a.x == b.x) { a.x = 0; b.x = 0; }
  if (a.y == b.y) { a.y = 0; b.y = 0; }
  if (a.z == b.z) { a.z = 0; b.z = 0; }
  Vec3<float> size = v3(fabsf(a.x) + fabsf(b.x), fabsf(a.y) + fabsf(b.y), fabsf(a.z) + fabsf(b.z));
  Vec3<float> mn = v3(xmin(vA.x, vB.x), xmin(vA.y, vB.y), xmin(vA.z, vB.z));
  Vec3<float> mx = v3(xmax(vA.x, vB.x), xmax(vA.y, vB.y), xmax(vA.z, vB.z));
  if (size.x > 0 && (p.x < mn.x || p.x > mx.x)) return false;
  if (size.y > 0 && (p.y < mn.y || p.y > mx.y)) return false;
  if (size.z > 0 && (p.z < mn.z || p.z > mx.z)) return false;
  t = tt;
  return true;
}

// ---------------------------------------------------------------------------
// Integrator-side primitive tests: hit in the open interval (tmin, tmax).
// ---------------------------------------------------------------------------
template <typename R>
NE_DEV bool isect_tri(const float4* tri, Vec3<R> o, Vec3<R> d, R tmin, R tmax, R& tout) {
  float4 a = tri[0], b = tri[1], c = tri[2];
  Vec3<R> v0 = v3f4<R>(a), e1 = v3f4<R>(b), e2 = v3f4<R>(c);
  Vec3<R> rov0 = o - v0;
  Vec3<R> n = cross(e1, e2), q = cross(rov0, d);
  R inv = R(1) / dot(d, n);
  R u = dot(-q, e2) * inv, v = dot(q, e1) * inv, t = dot(-n, rov0) * inv;
  if (!(t > tmin && t < tmax)) return false;
  if (u < R(0) || u > R(1) || v < R(0) || (u + v) > R(1)) return false;
  tout = t;
  return true;
}
template <typename R>
NE_DEV bool isect_sphere(float4 s, Vec3<R> o, Vec3<R> d, R tmin, R tmax, R& tout) {
  Vec3<R> c = v3f4<R>(s);
  R r = R(s.w);
  Vec3<R> oc = o - c;
  R a = dot(d, d), b = dot(oc, d), cc = dot(oc, oc) - r * r;
  R disc = b * b - a * cc;
  if (disc < R(0)) return false;
  R sq = xsqrt(disc);
  R t = (-b - sq) / a;
  if (!(t > tmin)) t = (-b + sq) / a;
  if (!(t > tmin && t < tmax)) return false;
  tout = t;
  return true;
}
template <typename R>
NE_DEV bool isect_rect(const float4* rc, Vec3<R> o, Vec3<R> d, R tmin, R tmax, R& tout) {
  float4 f0 = rc[0], f1 = rc[1], f2 = rc[2], f3 = rc[3];
  Vec3<R> p0 = v3f4<R>(f0), eu = v3f4<R>(f1), ev = v3f4<R>(f2), n = v3f4<R>(f3);
  R den = dot(n, d);
  if (xfabs(den) < R(float(kEPS))) return false;
  R t = dot(n, p0 - o) / den;
  if (!(t > tmin && t < tmax)) return false;
  Vec3<R> h = (o + t * d) - p0;
  R a = dot(h, eu) / dot(eu, eu), b = dot(h, ev) / dot(ev, ev);
  if (a < R(0) || a > R(1) || b < R(0) || b > R(1)) return false;
  tout = t;
  return true;
}

// Conservative slab test against a float box (Ize 2013: far distances scaled
// by 1 + 2 gamma(3)).  invd is finite (trav_init keeps |d| >= 2^-100 per
// axis), so no (plane - o) * invd is 0 * inf and min/max see no NaN.  The far
// scale is applied once to the smallest far distance: multiplying by a
// positive constant is monotone, so this equals scaling each one.
NE_DEV float fmin_(float a, float b) { return __builtin_fminf(a, b); }
NE_DEV double fmin_(double a, double b) { return __builtin_fmin(a, b); }
NE_DEV float fmax_(float a, float b) { return __builtin_fmaxf(a, b); }
NE_DEV double fmax_(double a, double b) { return __builtin_fmax(a, b); }
template <typename R>
NE_DEV bool slab(float mnx, float mny, float mnz, float mxx, float mxy, float mxz, Vec3<R> o, Vec3<R> invd, R tmax,
                 R& tent) {
  const R nx = (R(mnx) - o.x) * invd.x, fx = (R(mxx) - o.x) * invd.x;
  const R ny = (R(mny) - o.y) * invd.y, fy = (R(mxy) - o.y) * invd.y;
  const R nz = (R(mnz) - o.z) * invd.z, fz = (R(mxz) - o.z) * invd.z;
  const R t0 = fmax_(fmax_(R(0), fmin_(nx, fx)), fmax_(fmin_(ny, fy), fmin_(nz, fz)));
  const R far = fmin_(fmax_(nx, fx), fmin_(fmax_(ny, fy), fmax_(nz, fz))) * R(1.0000003576);
  tent = t0;
  return t0 <= fmin_(tmax, far);
}

template <typename R>
struct SurfHit {
  R t;
  int type;
  uint32_t key;   // original primitive index (ties break on (type, key))
  uint32_t slot;  // storage index
};


// Closest (or, with ANY, first) surface hit in (tmin, tmax).  Children are
// visited nearest first; the others go on the lane's stack.  Leaves are
// tested as soon as their box is hit.  Ties on t resolve to the lower
// (type, original index) key, as the oracle's linear scan does.
// Resumable BVH traversal: the per-lane state of one ray's closest-hit (or
// any-hit) query.  The current work item is either an inner node (its index)
// or a leaf (kItemLeaf | type << 29 | (count - 1) << 26 | first slot); one
// trav_step visits one node or tests one primitive, so kernels can interleave
// rays of different lengths in one wave (lane refill) and a wave never pays
// for a whole leaf's tests on every node step.  traverse() runs it to
// completion.  Closest-hit results do not depend on the visiting order: ties
// at equal t break on (type, original index).
template <typename R>
struct TravState {
  Vec3<R> o, d, invd;
  R tmin, tmax, best;
  uint32_t cur, leaf, bkey, bslot;  // leaf: parked leaf item (trav_step_wave), or kItemNone
  int sp, btype;
  bool found;
};

template <typename R>
NE_DEV void trav_init(TravState<R>& st, Vec3<R> o, Vec3<R> d, R tmin, R tmax) {
  st.o = o;
  st.d = d;
  // direction components below 2^-100 in magnitude count as +-2^-100 for the
  // box tests: invd stays finite, and the box intervals only widen
  const R tiny = R(7.888609052210118e-31);  // 2^-100
  auto nz = [tiny](R v) { return xfabs(v) < tiny ? (__builtin_signbit(v) ? -tiny : tiny) : v; };
  st.invd = v3(R(1) / nz(d.x), R(1) / nz(d.y), R(1) / nz(d.z));
  st.tmin = tmin;
  st.tmax = tmax;
  st.best = tmax;
  st.cur = 0;
  st.leaf = kItemNone;
  st.sp = 0;
  st.found = false;
  st.btype = 0;
  st.bkey = 0;
  st.bslot = 0;
}

// Byte c of a packed word of quantized planes, as a float.
NE_DEV float qbyte(uint32_t w, int c) { return float((w >> (8 * c)) & 0xFFu); }
// Child c of a wide node -> work item: an inner node index, or for a leaf
// (meta byte 0x80 | type << 4 | count) the packed leaf item.
NE_DEV uint32_t child_item(uint32_t ref, uint32_t meta, int c) {
  const uint32_t m = (meta >> (8 * c)) & 0xFFu;
  if (!(m & 0x80u)) return ref;
  return kItemLeaf | (((m >> 4) & 3u) << 29) | (((m & 15u) - 1u) << kItemCountShift) | ref;
}

template <typename R>
NE_DEV void trav_push(TravState<R>& st, uint32_t item, uint32_t* lds, int lane_stride, uint32_t* spill) {
  if (st.sp < kShortStack) lds[st.sp * lane_stride] = item;
  else spill[st.sp - kShortStack] = item;
  ++st.sp;
}

// Next item from the stack; true when the stack is empty (query complete).
template <typename R>
NE_DEV bool trav_pop(TravState<R>& st, const uint32_t* lds, int lane_stride, const uint32_t* spill) {
  if (st.sp == 0) return true;
  --st.sp;
  st.cur = st.sp < kShortStack ? lds[st.sp * lane_stride] : spill[st.sp - kShortStack];
  return false;
}

// Test of the first primitive of leaf item `it`: updates the closest hit and
// returns true on a hit for ANY.  next: the leaf's next primitive item, or
// kItemNone after its last.
template <typename R, bool ANY, bool C>
NE_DEV bool trav_prim(const DevScene& s, TravState<R>& st, uint32_t it, uint32_t& next, Count<C>& cnt) {
  const uint32_t ty = (it >> 29) & 3u, k = it & kItemFirstMask;
  R t;
  bool ok;
  uint32_t key = 0;
  cnt.prim();
  if (ty == PRIM_TRI) {
    const float4* tp = s.tris + 3 * size_t(k);
    ok = isect_tri<R>(tp, st.o, st.d, st.tmin, st.tmax, t);
    if (ok) key = __float_as_uint(tp[1].w);
  } else if (ty == PRIM_SPHERE) {
    ok = isect_sphere<R>(s.sph[k], st.o, st.d, st.tmin, st.tmax, t);
    if (ok) key = uint32_t(s.sph_orig[k]);
  } else {
    const float4* rc = s.rects + 4 * size_t(k);
    ok = isect_rect<R>(rc, st.o, st.d, st.tmin, st.tmax, t);
    if (ok) key = __float_as_uint(rc[2].w);
  }
  if (ok) {
    if (ANY) {
      st.found = true;
      return true;
    }
    if (t < st.best ||
        (st.found && t == st.best && (int(ty) < st.btype || (int(ty) == st.btype && key < st.bkey)))) {
      st.best = t; st.found = true; st.btype = int(ty); st.bkey = key; st.bslot = k;
    }
  }
  next = (it & (7u << kItemCountShift)) ? it + 1u - (1u << kItemCountShift) : kItemNone;
  return false;
}

// Visit of the inner node st.cur.  Returns true when nothing is left to visit
// (no child hit and the stack is empty).  Closest-hit queries go nearest child
// first; any-hit queries only move the nearest hit child to the front (one
// pass of three exchanges): any occluder ends the query.
template <typename R, bool ANY, bool C>
NE_DEV bool trav_node(const DevScene& s, TravState<R>& st, uint32_t* lds, int lane_stride, uint32_t* spill,
                      Count<C>& cnt) {
  const uint32_t it = st.cur;
  // 4-wide inner node (64 B): decode the four quantized child boxes, test
  // them, continue with the nearest hit child and push the others farthest
  // first (entry distances sorted by a five-exchange network; misses sort
  // last at +inf)
  const float4* nd = s.nodes + 4 * size_t(it);
  const float4 f0 = nd[0];
  const uint4 f1 = reinterpret_cast<const uint4*>(nd)[1], f2 = reinterpret_cast<const uint4*>(nd)[2];
  const int4 ch = reinterpret_cast<const int4*>(nd)[3];
  cnt.node();
  const uint32_t ex = __float_as_uint(f0.w);
  const float sx = __uint_as_float((ex & 0xFFu) << 23), sy = __uint_as_float(((ex >> 8) & 0xFFu) << 23),
              sz = __uint_as_float(((ex >> 16) & 0xFFu) << 23);
  const R inf = xinf<R>();
  R t0 = inf, t1 = inf, t2 = inf, t3 = inf, te;
#define NE_CHILD_SLAB(c, ref, tout)                                                                           \
  if ((ref) >= 0 &&                                                                                         \
      slab<R>(__builtin_fmaf(qbyte(f1.x, c), sx, f0.x), __builtin_fmaf(qbyte(f1.y, c), sy, f0.y),             \
              __builtin_fmaf(qbyte(f1.z, c), sz, f0.z), __builtin_fmaf(qbyte(f1.w, c), sx, f0.x),             \
              __builtin_fmaf(qbyte(f2.x, c), sy, f0.y), __builtin_fmaf(qbyte(f2.y, c), sz, f0.z), st.o,       \
              st.invd, st.best, te))                                                                        \
    tout = te;
  NE_CHILD_SLAB(0, ch.x, t0)
  NE_CHILD_SLAB(1, ch.y, t1)
  NE_CHILD_SLAB(2, ch.z, t2)
  NE_CHILD_SLAB(3, ch.w, t3)
#undef NE_CHILD_SLAB
  uint32_t i0 = child_item(uint32_t(ch.x), f2.z, 0);
  uint32_t i1 = child_item(uint32_t(ch.y), f2.z, 1);
  uint32_t i2 = child_item(uint32_t(ch.z), f2.z, 2);
  uint32_t i3 = child_item(uint32_t(ch.w), f2.z, 3);
  auto order = [](R& ta, uint32_t& ia, R& tb, uint32_t& ib) {
    if (tb < ta) {
      const R t = ta; ta = tb; tb = t;
      const uint32_t i = ia; ia = ib; ib = i;
    }
  };
  if constexpr (ANY) {
    order(t2, i2, t3, i3);
    order(t1, i1, t2, i2);
    order(t0, i0, t1, i1);
  } else {
    order(t0, i0, t1, i1);
    order(t2, i2, t3, i3);
    order(t0, i0, t2, i2);
    order(t1, i1, t3, i3);
    order(t1, i1, t2, i2);
  }
  if (!(t0 < inf)) return trav_pop(st, lds, lane_stride, spill);
  if (t3 < inf) trav_push(st, i3, lds, lane_stride, spill);
  if (t2 < inf) trav_push(st, i2, lds, lane_stride, spill);
  if (t1 < inf) trav_push(st, i1, lds, lane_stride, spill);
  st.cur = i0;
  return false;
}

// One node visit or one primitive test.  Returns true when the query is
// complete: the stack is empty, or (ANY) a hit was found.  lds: this lane's
// short stack (entries lane_stride apart); spill: its private overflow array.
template <typename R, bool ANY, bool C>
NE_DEV bool trav_step(const DevScene& s, TravState<R>& st, uint32_t* lds, int lane_stride, uint32_t* spill,
                      Count<C>& cnt) {
  const uint32_t it = st.cur;
  if (it & kItemLeaf) {
    uint32_t next;
    if (trav_prim<R, ANY, C>(s, st, it, next, cnt)) return true;
    if (next != kItemNone) {  // more primitives in this leaf
      st.cur = next;
      return false;
    }
    return trav_pop(st, lds, lane_stride, spill);
  }
  return trav_node<R, ANY, C>(s, st, lds, lane_stride, spill, cnt);
}

// Speculative traversal for the lane-refill kernels (postponed leaves, Aila
// and Laine 2009).  Each lane parks up to one leaf in st.leaf and goes on
// visiting inner nodes past it.  Per call the wave votes: when at most half
// of its lanes stand at an inner node, every lane with a parked leaf tests one
// of its primitives; otherwise the lanes at inner nodes visit one node each.
// The node and primitive paths thus each run with most lanes busy, where a
// per-lane step runs both on nearly every call.  Visiting order does not
// change the closest hit (ties break on type and original index); parked
// leaves only delay the shrinking of the search interval.  Every lane with a
// live query calls it; returns true when this lane's query is complete.
template <typename R, bool ANY, bool C>
NE_DEV bool trav_step_wave(const DevScene& s, TravState<R>& st, uint32_t* lds, int lane_stride, uint32_t* spill,
                           Count<C>& cnt) {
  const bool at_node = !(st.cur & kItemLeaf);
  const unsigned long long busy = __ballot(true), nodes = __ballot(at_node);
  if (2 * __popcll(nodes) <= __popcll(busy)) {
    if (st.leaf != kItemNone) {
      uint32_t next;
      if (trav_prim<R, ANY, C>(s, st, st.leaf, next, cnt)) return true;
      st.leaf = next;
    }
  } else if (at_node) {
    if (trav_node<R, ANY, C>(s, st, lds, lane_stride, spill, cnt)) st.cur = kItemNone;
  }
  if (st.leaf == kItemNone && st.cur != kItemNone && (st.cur & kItemLeaf)) {  // park the leaf in front
    st.leaf = st.cur;
    if (trav_pop(st, lds, lane_stride, spill)) st.cur = kItemNone;
  }
  return st.cur == kItemNone && st.leaf == kItemNone;
}

template <typename R, bool ANY, bool C>
NE_DEV bool traverse(const DevScene& s, Vec3<R> o, Vec3<R> d, R tmin, R tmax, uint32_t* lds, int lane_stride,
                     SurfHit<R>& hit, Count<C>& cnt) {
  if (s.n_nodes == 0) return false;
  uint32_t spill[kSpillStack];
  TravState<R> st;
  trav_init(st, o, d, tmin, tmax);
  while (!trav_step<R, ANY, C>(s, st, lds, lane_stride, spill, cnt)) {
  }
  if (!st.found) return false;
  hit.t = st.best; hit.type = st.btype; hit.key = st.bkey; hit.slot = st.bslot;
  return true;
}

// Geometric normal and material of a surface hit.
template <typename R>
NE_DEV void hit_normal(const DevScene& s, const SurfHit<R>& h, Vec3<R> o, Vec3<R> d, Vec3<R>& n, int& mat) {
  if (h.type == PRIM_TRI) {
    const float4* tp = s.tris + 3 * size_t(h.slot);
    float4 a = tp[0];
    n = normalize(cross(v3f4<R>(tp[1]), v3f4<R>(tp[2])));
    mat = __float_as_int(a.w);
  } else if (h.type == PRIM_SPHERE) {
    float4 sp = s.sph[h.slot];
    n = ((o + h.t * d) - v3f4<R>(sp)) / R(sp.w);
    mat = s.sph_mat[h.slot];
  } else {
    const float4* rc = s.rects + 4 * size_t(h.slot);
    n = v3f4<R>(rc[3]);
    mat = __float_as_int(rc[0].w);
  }
}

// Material of the surface hit coded as type << 30 | storage slot.
NE_DEV int hit_material(const DevScene& s, uint32_t ref) {
  const int type = int(ref >> 30);
  const uint32_t slot = ref & 0x3FFFFFFFu;
  if (type == PRIM_TRI) return __float_as_int(s.tris[3 * size_t(slot)].w);
  if (type == PRIM_SPHERE) return s.sph_mat[slot];
  return __float_as_int(s.rects[4 * size_t(slot)].w);
}

// ---------------------------------------------------------------------------
// Volume: unit-cube interval, trilinear density, majorant DDA, tracking.
// ---------------------------------------------------------------------------
template <typename R>
NE_DEV Vec3<R> xf_point(const float* m, Vec3<R> p) {
  return v3(R(m[0]) * p.x + R(m[1]) * p.y + R(m[2]) * p.z + R(m[3]),
            R(m[4]) * p.x + R(m[5]) * p.y + R(m[6]) * p.z + R(m[7]),
            R(m[8]) * p.x + R(m[9]) * p.y + R(m[10]) * p.z + R(m[11]));
}
template <typename R>
NE_DEV Vec3<R> xf_vec(const float* m, Vec3<R> p) {
  return v3(R(m[0]) * p.x + R(m[1]) * p.y + R(m[2]) * p.z, R(m[4]) * p.x + R(m[5]) * p.y + R(m[6]) * p.z,
            R(m[8]) * p.x + R(m[9]) * p.y + R(m[10]) * p.z);
}

template <typename R>
NE_DEV bool box_interval(const DevScene& s, Vec3<R> oo, Vec3<R> dd, R& ta, R& tb) {
  R t0 = -xinf<R>(), t1 = xinf<R>();
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    R oa = comp(oo, a), da = comp(dd, a);
    R id = R(1) / da;
    R tn = (R(-0.5) - oa) * id, tf = (R(0.5) - oa) * id;
    if (tn > tf) { R x = tn; tn = tf; tf = x; }
    t0 = tn > t0 ? tn : t0;
    t1 = tf < t1 ? tf : t1;
  }
  ta = t0; tb = t1;
  return t0 <= t1;
}

NE_DEV float grid_at(const DevScene& s, int x, int y, int z) {
  return s.grid[grid_offset(uint32_t(x), uint32_t(y), uint32_t(z), s.grid_step_y, s.grid_step_z)];
}
// GridMedia::sampleAt -> interpolatedDensity (src/materials/GridMedia.cpp:25-43):
// OCS [-.5,.5]^3 -> grid space (p + .5) * res, clamp, cell-corner trilinear,
// corners at index >= res read as 0 (the grid's zero padding).
// A homogeneous medium (HomogeneousMedia, src/materials/HomogeneousMedia.cpp)
// reads density 1 everywhere; the multiplier carries its density.
// SMALL (grid_small scenes): corner addresses are 32-bit byte offsets from
// the grid base (uniform), so each load takes a scalar base and a 32-bit lane
// offset and no lane does 64-bit address arithmetic; same corners, same value.
template <bool SMALL = false>
NE_DEV float density_gcs(const DevScene& s, Vec3<float> gp) {
  if (s.homog) return 1.0f;
  // clamp to [0, res] (std::max with 0, then std::min with res): for finite
  // values the median of (x, 0, res) is that clamp, one v_med3_f32
  gp = v3(__builtin_amdgcn_fmed3f(gp.x, 0.0f, float(s.res[0])), __builtin_amdgcn_fmed3f(gp.y, 0.0f, float(s.res[1])),
          __builtin_amdgcn_fmed3f(gp.z, 0.0f, float(s.res[2])));
  if constexpr (SMALL) {
    // x >= 0 here: the truncating conversion is floor(x), and v_fract_f32
    // returns x - floor(x), which is exact (the low bits of x): the same
    // index and fraction as int(floorf(x)) and x - float(ix) below
    const uint32_t ix = uint32_t(gp.x), iy = uint32_t(gp.y), iz = uint32_t(gp.z);
    const float dx = __builtin_amdgcn_fractf(gp.x), dy = __builtin_amdgcn_fractf(gp.y),
                dz = __builtin_amdgcn_fractf(gp.z);
    // both steps are under 2^24 and every offset under 2^32: 24-bit multiplies are exact
    const uint32_t b = (__umul24(iz >> 1, uint32_t(s.grid_step_z)) + __umul24(iy >> 2, uint32_t(s.grid_step_y)) +
                        (ix >> 2) * kBrickFloats +
                        (((iz & 1u) << 4) | ((iy & 3u) << 2) | (ix & 3u))) * 4u;
    const uint32_t bx = (ix & 3u) != 3u ? 4u : (kBrickFloats - 3u) * 4u;  // x + 1: same brick or the next
    const uint32_t by = (iy & 3u) != 3u ? 16u : (s.grid_step_y - 12u) * 4u;
    const uint32_t bz = (iz & 1u) == 0u ? 64u : (s.grid_step_z - 16u) * 4u;
    const char* g = reinterpret_cast<const char*>(s.grid);
    auto at = [&](uint32_t o) { return *reinterpret_cast<const float*>(g + size_t(o)); };
    const float c000 = at(b), c100 = at(b + bx), c010 = at(b + by), c110 = at(b + by + bx);
    const float c001 = at(b + bz), c101 = at(b + bz + bx), c011 = at(b + by + bz), c111 = at(b + by + bz + bx);
    const float d00 = mixf(c000, c100, dx), d10 = mixf(c010, c110, dx);
    const float d01 = mixf(c001, c101, dx), d11 = mixf(c011, c111, dx);
    const float d0 = mixf(d00, d10, dy), d1 = mixf(d01, d11, dy);
    return mixf(d0, d1, dz);
  }
  int ix = int(floorf(gp.x)), iy = int(floorf(gp.y)), iz = int(floorf(gp.z));
  float dx = gp.x - float(ix), dy = gp.y - float(iy), dz = gp.z - float(iz);
  // 0 <= i <= res on every axis and the padding holds index res + 1: every
  // corner is in the allocation.  Within a brick the x-neighbour is the next
  // float, so each (y, z) corner row is one 8-byte read (4-byte aligned); a
  // corner row that crosses into the next brick in x (ix % 4 == 3) takes its
  // second float from there (the first read's second float then lies in the
  // same brick or the next one, which the allocation's slack covers).
  const float* p = s.grid + grid_offset(uint32_t(ix), uint32_t(iy), uint32_t(iz), s.grid_step_y, s.grid_step_z);
  const uint32_t oy = (iy & 3) != 3 ? 4u : s.grid_step_y - 12u;
  const uint32_t oz = (iz & 1) == 0 ? 16u : s.grid_step_z - 16u;
  float2 q00, q10, q01, q11;
  __builtin_memcpy(&q00, p, 8);
  __builtin_memcpy(&q10, p + oy, 8);
  __builtin_memcpy(&q01, p + oz, 8);
  __builtin_memcpy(&q11, p + oy + oz, 8);
  if ((ix & 3) == 3) {
    constexpr uint32_t ox = kBrickFloats - 3u;  // x + 1 in the next brick
    q00.y = p[ox];
    q10.y = p[oy + ox];
    q01.y = p[oz + ox];
    q11.y = p[oy + oz + ox];
  }
  const float c000 = q00.x, c100 = q00.y, c010 = q10.x, c110 = q10.y;
  const float c001 = q01.x, c101 = q01.y, c011 = q11.x, c111 = q11.y;
  float d00 = mixf(c000, c100, dx), d10 = mixf(c010, c110, dx);
  float d01 = mixf(c001, c101, dx), d11 = mixf(c011, c111, dx);
  float d0 = mixf(d00, d10, dy), d1 = mixf(d01, d11, dy);
  return mixf(d0, d1, dz);
}
template <bool SMALL = false>
NE_DEV float density_ocs(const DevScene& s, Vec3<float> p) {
  return density_gcs<SMALL>(s, (p + v3s(0.5f)) * v3(float(s.res[0]), float(s.res[1]), float(s.res[2])));
}

// 1 / mu for the tracking step: in FP32 the cell's stored reciprocal (the
// correctly rounded quotient, the same bits as the division), in FP64 the
// division itself.
NE_DEV float maj_inverse(float, float inv) { return inv; }
NE_DEV double maj_inverse(double mu, float) { return 1.0 / mu; }

// log(1 - u) of a draw u in [0, 1 - 2^-24].  The argument is a normal float
// in [2^-24, 1], for which the device logf is log2 by v_log_f32 times ln 2
// carried in two parts (0x1.62e42ep-1 + 0x1.efa39ep-25, the product's error
// kept by an fma); these are those operations without its subnormal scaling
// and its infinity and NaN selects, so the result has the same bits.
NE_DEV float log1m_draw(float u) {
  const float y = __builtin_amdgcn_logf(1.0f - u);
  const float hi = y * 0x1.62e42ep-1f;
  float e = __builtin_fmaf(y, 0x1.62e42ep-1f, -hi);
  e = __builtin_fmaf(y, 0x1.efa39ep-25f, e);
  return hi + e;
}
NE_DEV double log1m_draw(double u) { return xlog(1.0 - u); }

// 3-D DDA over the majorant grid along the OCS ray (oo, dd) for t in [ta, tb).
// Calls f(c0, c1, mu, inv) per cell with mu > 0 (inv its stored FP32
// reciprocal); f returns false to stop.
// Empty cells may hold -D instead of 0, D >= 1 the L1 (Manhattan) distance
// to the nearest non-empty cell: each step moves one cell along one axis, so
// the next D - 1 cells are empty too and their majorants are not read.  Cell
// boundaries come from the cell index, never accumulated, so the t of every
// non-empty cell is the same with or without skipping.
template <typename R, bool C, typename F>
NE_DEV void majorant_dda(const DevScene& s, Vec3<R> oo, Vec3<R> dd, R ta, R tb, Count<C>& cnt, F&& f) {
  const int M = s.mres;
  const R Mr = R(M), invM = R(1) / Mr;
  int cell[3], step[3];
  R tnext[3], inv[3];
  // clip [ta, tb) to the box of the cells with a positive majorant: the box
  // faces are cell boundaries in the DDA's own arithmetic, so the cells and
  // boundary distances inside are those of a walk from the volume box
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const R da = comp(dd, a), oa = comp(oo, a);
    inv[a] = da != R(0) ? R(1) / da : R(0);
    const R blo = R(s.occ_lo[a]) * invM - R(0.5), bhi = R(s.occ_hi[a] + 1) * invM - R(0.5);
    if (s.occ_lo[a] > s.occ_hi[a]) {
      tb = ta;  // no cell has a positive majorant
    } else if (da != R(0)) {
      const R t0 = (blo - oa) * inv[a], t1 = (bhi - oa) * inv[a];
      ta = xmax(ta, xmin(t0, t1));
      tb = xmin(tb, xmax(t0, t1));
    } else if (!(oa >= blo && oa <= bhi)) {
      tb = ta;  // parallel to the box and outside it
    }
  }
  if (!(ta < tb)) return;
  Vec3<R> pe = oo + ta * dd;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    R g = (comp(pe, a) + R(0.5)) * Mr;
    int c = int(xfloor(g));
    c = c < 0 ? 0 : (c > M - 1 ? M - 1 : c);
    cell[a] = c;
    R da = comp(dd, a), oa = comp(oo, a);
    if (da > R(0)) { step[a] = 1; tnext[a] = ((R(c + 1) * invM - R(0.5)) - oa) * inv[a]; }
    else if (da < R(0)) { step[a] = -1; tnext[a] = ((R(c) * invM - R(0.5)) - oa) * inv[a]; }
    else { step[a] = 0; tnext[a] = xinf<R>(); }
  }
  R t = ta;
  int skip = 0;
  while (t < tb) {
    // exit axis: 1 if tnext[1] < tnext[0], then 2 if tnext[2] is below that
    // one (ties keep the lower axis); kept as the two comparisons, which the
    // selects below use as they are
    const bool s1 = tnext[1] < tnext[0];
    const R t01 = s1 ? tnext[1] : tnext[0];
    const bool s2 = tnext[2] < t01;
    const R tna = s2 ? tnext[2] : t01;
    R tc = xmin(tna, tb);
    cnt.cell();
    if (skip > 0) {
      --skip;
    } else {
      // M <= 256: 24-bit multiplies give the exact cell index, and the
      // table (at most 2^24 pairs) is read by a 32-bit byte offset
      const uint32_t ci = uint32_t(cell[0]) + __umul24(uint32_t(cell[1]), uint32_t(M)) +
                          __umul24(uint32_t(cell[2]), uint32_t(M) * uint32_t(M));
      const float2 mp =
          *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(s.maj_pair) + size_t(ci * 8u));
      const float raw = mp.x;
      if (raw < 0.0f) skip = int(-raw) - 1;
      else if (raw > 0.0f && tc > t && !f(t, tc, R(raw), mp.y)) return;
    }
    t = tc;
    if (tna >= tb) break;
    const int sa = s2 ? step[2] : (s1 ? step[1] : step[0]);
    const int c = (s2 ? cell[2] : (s1 ? cell[1] : cell[0])) + sa;
    if (c < 0 || c >= M) break;
    const R bnd = sa > 0 ? R(c + 1) * invM : R(c) * invM;
    const R oa = s2 ? oo.z : (s1 ? oo.y : oo.x);
    const R nt = ((bnd - R(0.5)) - oa) * (s2 ? inv[2] : (s1 ? inv[1] : inv[0]));
    const bool a1 = s1 && !s2, a0 = !s1 && !s2;
    cell[0] = a0 ? c : cell[0];
    tnext[0] = a0 ? nt : tnext[0];
    cell[1] = a1 ? c : cell[1];
    tnext[1] = a1 ? nt : tnext[1];
    cell[2] = s2 ? c : cell[2];
    tnext[2] = s2 ? nt : tnext[2];
  }
}

template <typename R, bool C, bool SMALL = false>
NE_DEV R sigma_at(const DevScene& s, Vec3<R> oo, Vec3<R> dd, R t, Count<C>& cnt) {
  Vec3<R> p = oo + t * dd;
  cnt.lookup();
  return R(density_ocs<SMALL>(s, v3(float(p.x), float(p.y), float(p.z)))) * R(s.sigma_bar);
}

// Delta tracking (GridMedia::sample, src/materials/GridMedia.cpp:71-98) with
// per-cell majorants: step t -= log(1 - xi) * (1 / mu), accept if
// sigma * (1 / mu) > xi', the reference's multiply-by-inverse form.
template <typename R, bool C, bool SMALL = false>
NE_DEV bool delta_track(const DevScene& s, Vec3<R> o, Vec3<R> d, R ta, R tb, PathRng& rng, Count<C>& cnt, R& th) {
  Vec3<R> oo = xf_point<R>(s.inv, o), dd = xf_vec<R>(s.inv, d);
  bool hit = false;
  majorant_dda<R, C>(s, oo, dd, ta, tb, cnt, [&](R c0, R c1, R mu, float inv) {
    if (!(mu > R(0)) || xisinf(mu)) return true;  // empty cell; grids are validated finite at upload
    const R imu = maj_inverse(mu, inv);
    R t = c0;
    while (true) {
      R r = R(rng.next_keep1());
      t -= log1m_draw(r) * imu;
      if (!(t < c1)) return true;
      R sig = sigma_at<R, C, SMALL>(s, oo, dd, t, cnt);
      R ra = R(rng.next_buffered());
      if (sig * imu > ra) { hit = true; th = t; return false; }
    }
  });
  return hit;
}

// Ratio tracking with the reference's Russian roulette (GridMedia.cpp:54-66).
template <typename R, bool C, bool SMALL = false>
// A homogeneous medium takes the expectation in closed form,
// HomogeneousMedia::Tr (HomogeneousMedia.cpp:15-18) with the grey extinction
// of the tracking step law, and draws nothing.
NE_DEV R ratio_track(const DevScene& s, Vec3<R> o, Vec3<R> d, R ta, R tb, PathRng& rng, Count<C>& cnt) {
  if (s.homog) return xexp(-R(s.sigma_bar) * (tb - ta));
  Vec3<R> oo = xf_point<R>(s.inv, o), dd = xf_vec<R>(s.inv, d);
  R Tr = R(1);
  majorant_dda<R, C>(s, oo, dd, ta, tb, cnt, [&](R c0, R c1, R mu, float inv) {
    if (!(mu > R(0)) || xisinf(mu)) return true;  // empty cell; grids are validated finite at upload
    const R imu = maj_inverse(mu, inv);
    R t = c0;
    while (true) {
      t -= log1m_draw(R(rng.next_keep1())) * imu;
      if (!(t < c1)) return true;
      R sig = sigma_at<R, C, SMALL>(s, oo, dd, t, cnt);
      Tr *= R(1) - xmax(R(0), sig * imu);
      if (Tr < R(0.1)) {
        R q = xmax(R(0.05), R(1) - Tr);
        if (R(rng.next_buffered()) < q) { Tr = R(0); return false; }
        Tr /= R(1) - q;
      }
    }
  });
  return Tr;
}

// ---------------------------------------------------------------------------
// Phase function and surface BSDF
// ---------------------------------------------------------------------------
// HG::eval, src/materials/Medium.h:116-128 (reference argument convention)
NE_DEV float ref_hg_eval(float g, Vec3<float> a, Vec3<float> b) {
  float c = dot(normalize(a), normalize(b));
  float den = 1.0f + g * g - 2.0f * g * c;
  return kINV4PI * (1.0f - g * g) / (den * sqrtf(den));
}
// HG::sample local frame, src/materials/Medium.h:96-114
NE_DEV Vec3<float> ref_hg_sample(float g, float u0, float u1) {
  float ct;
  if (fabsf(g) < 1e-3f) ct = 1.0f - 2.0f * u0;
  else {
    float sqr = (1.0f - g * g) / (1.0f + g - 2.0f * g * u0);
    sqr = sqr * sqr;
    ct = -1.0f / (2.0f * g) * (1.0f + g * g - sqr);
  }
  float st = sqrtf(1.0f - ct * ct);
  float ph = float(2.0f * kPI * u1);
  return v3(cosf(ph) * st, sinf(ph) * st, ct);
}
// HG in the propagation convention: p(cos) with cos = dot(d, wi).
template <typename R>
NE_DEV R hg_p(R g, R c) {
  R den = R(1) + g * g - R(2) * g * c;
  return R(kINV4PI) * (R(1) - g * g) / (den * xsqrt(den));
}
// Inverse-CDF sample of the same density, framed on d (sign flipped
// relative to the reference sampler so g > 0 scatters forward).
template <typename R>
NE_DEV Vec3<R> hg_sample(R g, Vec3<R> d, R u0, R u1) {
  R ct;
  if (xfabs(g) < R(1e-3)) ct = R(1) - R(2) * u0;
  else {
    R sq = (R(1) - g * g) / (R(1) + g - R(2) * g * u0);
    ct = (R(1) + g * g - sq * sq) / (R(2) * g);
  }
  ct = xmin(R(1), xmax(R(-1), ct));
  R st = xsqrt(xmax(R(0), R(1) - ct * ct));
  R ph = R(kTWO_PI) * u1, sp, cp;
  xsincos(ph, sp, cp);
  Vec3<R> ss, ts;
  onb(d, ss, ts);
  return to_world(v3(st * cp, st * sp, ct), d, ss, ts);
}

// GGXDistribution::D and ::pdf, src/core/Microfacet.cpp:5-15,58-65
NE_DEV float ref_ggx_D(float alpha, Vec3<float> h) {
  float a2 = alpha * alpha, c = h.z, c2 = c * c;
  float den = c2 * (a2 - 1.0f) + 1.0f;
  den = xmax(float(kEPS), float(kPI * den * den));
  return a2 / den;
}
NE_DEV float ref_ggx_pdf(float alpha, Vec3<float> sc, Vec3<float> h) {
  return ref_ggx_D(alpha, h) * h.z / (4 * dot(sc, h));
}

// Glossy lobe of GlossyBSDF (src/core/GlossyBSDF.cpp:26-48: GGX D, Schlick-GGX
// G with k = alpha/2, Schlick Fresnel f0 = 0.04, Lambert term weighted by
// (1-F)(1-metallic)), with the cosine applied by the caller, G on the shading
// normal, and sampled as a mixture of GGX half-vector sampling
// (reflection about h) and cosine-weighted hemisphere sampling.
template <typename R>
struct Glossy {
  Vec3<R> albedo;
  R alpha, metallic;
  NE_DEV static R D(R alpha, R ch) {
    R a2 = alpha * alpha, c2 = ch * ch;
    R den = c2 * (a2 - R(1)) + R(1);
    den = xmax(R(kEPS), R(kPI) * den * den);
    return a2 / den;
  }
  NE_DEV static R G1(R alpha, R c) {
    R k = alpha / R(2);
    return c / xmax(R(kEPS), c * (R(1) - k) + k);
  }
  NE_DEV static R F(R c) {
    R m = R(1) - c, m2 = m * m;
    return R(0.04) + R(0.96) * (m2 * m2 * m);
  }
  NE_DEV R p_spec(R co) const {
    R fs = F(co);
    R kd = avg3(albedo) * (R(1) - fs) * (R(1) - metallic);
    R p = fs / (fs + kd);
    return xmin(R(0.9), xmax(R(0.1), p));
  }
  NE_DEV Vec3<R> eval(Vec3<R> wo, Vec3<R> wi) const {
    if (wo.z <= R(0) || wi.z <= R(0)) return v3s(R(0));
    Vec3<R> h = normalize(wo + wi);
    R Fh = F(dot(wo, h));
    R spec = D(alpha, h.z) * G1(alpha, wo.z) * G1(alpha, wi.z) * Fh / (R(4) * wo.z * wi.z);
    R kd = (R(1) - Fh) * (R(1) - metallic);
    return albedo * (kd / R(kPI)) + v3s(spec);
  }
  NE_DEV R pdf(Vec3<R> wo, Vec3<R> wi) const {
    if (wo.z <= R(0) || wi.z <= R(0)) return R(0);
    Vec3<R> h = normalize(wo + wi);
    R ps = p_spec(wo.z);
    R pspec = D(alpha, h.z) * h.z / (R(4) * xfabs(dot(wo, h)));
    R pdiff = wi.z / R(kPI);
    return ps * pspec + (R(1) - ps) * pdiff;
  }
  // eval() and pdf() of one direction pair with the half vector, D, the
  // Fresnel cosine and the lobe choice computed once (the same values)
  NE_DEV void eval_pdf(Vec3<R> wo, Vec3<R> wi, Vec3<R>& f, R& p) const {
    if (wo.z <= R(0) || wi.z <= R(0)) {
      f = v3s(R(0));
      p = R(0);
      return;
    }
    Vec3<R> h = normalize(wo + wi);
    const R dh = D(alpha, h.z), oh = dot(wo, h);
    R Fh = F(oh);
    R spec = dh * G1(alpha, wo.z) * G1(alpha, wi.z) * Fh / (R(4) * wo.z * wi.z);
    R kd = (R(1) - Fh) * (R(1) - metallic);
    f = albedo * (kd / R(kPI)) + v3s(spec);
    R ps = p_spec(wo.z);
    R pspec = dh * h.z / (R(4) * xfabs(oh));
    R pdiff = wi.z / R(kPI);
    p = ps * pspec + (R(1) - ps) * pdiff;
  }
  NE_DEV Vec3<R> sample(Vec3<R> wo, R u0, R u1, R u2) const {
    R ps = p_spec(wo.z);
    if (u0 < ps) {
      R a2 = alpha * alpha;
      R ct = xmin(R(1), xsqrt((R(1) - u1) / xmax(R(kEPS), u1 * (a2 - R(1)) + R(1))));
      R st = xsqrt(xmax(R(0), R(1) - ct * ct));
      R ph = R(kTWO_PI) * u2, sp, cp;
      xsincos(ph, sp, cp);
      Vec3<R> h = v3(st * cp, st * sp, ct);
      return R(2) * dot(wo, h) * h - wo;
    }
    R r = xsqrt(u1), ph = R(kTWO_PI) * u2, sp, cp;
    xsincos(ph, sp, cp);
    return v3(r * cp, r * sp, xsqrt(xmax(R(0), R(1) - u1)));
  }
};

// ---------------------------------------------------------------------------
// Lights
// ---------------------------------------------------------------------------
// Environment map (InfiniteAreaLight, src/lights/InfiniteAreaLight.h:8-130):
// latitude-longitude RGB texels looked up as InfiniteAreaLight::Le does, in
// the frame generateOrthonormalCS builds around (0,0,1): local
// (x, y, z) = (d.y, -d.x, d.z), u = phi / 2 pi with phi = atan2(y, x) in
// [0, 2 pi), v = theta / pi with theta = acos(z); row 0 is theta = 0.
// Texels are piecewise constant.  Sampling inverts a 2-D piecewise-constant
// distribution over luminance * sin(theta) (Distribution2D,
// src/utils/Sampling.h:76-112, with each row's conditional over the row's W
// texels) and uses the same frame as the lookup, so that the solid-angle pdf
// p(u, v) / (2 pi^2 sin theta) belongs to the radiance it weights.
template <typename R>
NE_DEV int env_texel(const DevScene& s, Vec3<R> d, R& sin_t) {
  R phi = xatan2(-d.x, d.y);
  phi = phi < R(0) ? phi + R(kTWO_PI) : phi;
  const R z = xmin(R(1), xmax(R(-1), d.z));
  sin_t = xsqrt(xmax(R(0), R(1) - z * z));
  const R theta = xacos(z);
  int i = int(phi * R(1.0 / kTWO_PI) * R(s.env_w));
  int j = int(theta * R(1.0 / kPI) * R(s.env_h));
  i = i < s.env_w - 1 ? (i > 0 ? i : 0) : s.env_w - 1;  // also keeps a NaN direction in range
  j = j < s.env_h - 1 ? (j > 0 ? j : 0) : s.env_h - 1;
  return j * s.env_w + i;
}
// Largest k in [0, n) with cdf[k] <= u (cdf[0] = 0, cdf[n] = 1, u in [0, 1)).
NE_DEV int cdf_find(const float* cdf, int n, float u) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (cdf[mid] <= u) lo = mid;
    else hi = mid;
  }
  return lo;
}
template <typename R>
NE_DEV R env_map_pdf(const DevScene& s, Vec3<R> d) {
  R st;
  const int k = env_texel<R>(s, d, st);
  if (!(st > R(0))) return R(0);
  return R(s.env_func[k]) * R(s.env_inv_int) / (R(2.0 * kPI * kPI) * st);
}

template <typename R>
NE_DEV Vec3<R> env_le(const DevScene& s, Vec3<R> d) {
  if (s.env_type == 3) {
    R st;
    const float* t = s.env_tex + 3 * size_t(env_texel<R>(s, d, st));
    return v3(R(t[0]), R(t[1]), R(t[2]));
  }
  if (s.env_type == 1) return v3(R(s.env[0]), R(s.env[1]), R(s.env[2]));
  if (s.env_type == 2) {
    if (d.y >= R(0))
      return v3(R(s.env[3]), R(s.env[4]), R(s.env[5])) * (R(1) - d.y) + v3(R(s.env[0]), R(s.env[1]), R(s.env[2])) * d.y;
    return v3(R(s.env[6]), R(s.env[7]), R(s.env[8]));
  }
  return v3s(R(0));
}
// Solid-angle pdf of env_sample for direction d.
template <typename R>
NE_DEV R env_pdf(const DevScene& s, Vec3<R> d) {
  return s.env_type == 3 ? env_map_pdf<R>(s, d) : R(1.0 / kFOUR_PI);
}
// A direction toward the environment: uniform on the sphere for the constant
// and sky environments, importance-sampled for a map.
template <typename R>
NE_DEV void env_sample(const DevScene& s, R ua, R ub, Vec3<R>& wi, R& pl, Vec3<R>& Le) {
  if (s.env_type != 3) {
    R z = R(1) - R(2) * ua, r = xsqrt(xmax(R(0), R(1) - z * z)), ph = R(kTWO_PI) * ub, sp, cp;
    xsincos(ph, sp, cp);
    wi = v3(r * cp, r * sp, z);
    Le = env_le<R>(s, wi);
    pl = R(1.0 / kFOUR_PI);
    return;
  }
  const int W = s.env_w, H = s.env_h;
  const float fv = float(ub), fu = float(ua);
  const int j = cdf_find(s.env_marg, H, fv);
  const float m0 = s.env_marg[j], m1 = s.env_marg[j + 1];
  const float dv = m1 > m0 ? (fv - m0) / (m1 - m0) : 0.0f;
  const float* row = s.env_cdf + size_t(j) * size_t(W + 1);
  const int i = cdf_find(row, W, fu);
  const float c0 = row[i], c1 = row[i + 1];
  const float du = c1 > c0 ? (fu - c0) / (c1 - c0) : 0.0f;
  const R u = (R(i) + R(du)) / R(W), v = (R(j) + R(dv)) / R(H);
  const R theta = v * R(kPI), phi = u * R(kTWO_PI);
  R st, ct, sp, cp;
  xsincos(theta, st, ct);
  xsincos(phi, sp, cp);
  wi = v3(-st * sp, st * cp, ct);
  const size_t k = size_t(j) * size_t(W) + size_t(i);
  pl = st > R(0) ? R(s.env_func[k]) * R(s.env_inv_int) / (R(2.0 * kPI * kPI) * st) : R(0);
  const float* t = s.env_tex + 3 * k;
  Le = v3(R(t[0]), R(t[1]), R(t[2]));
}

// Emissive primitives: rectangles, spheres and triangles with an emitter
// material (DiffuseLight on a primitive, src/lights/DiffuseLight.cpp:8-20).
// Light codes are type << 30 | storage slot.  Emission is two-sided.
//   rectangle  uniform point on the quad, area pdf -> solid angle
//              (Rectangle::samplePointOnSurface/pdf, src/primitives/Rectangle.cpp:78-88,159-171)
//   triangle   uniform point (b1 = v sqrt(u), b2 = sqrt(u)(1 - v)), area pdf
//              (Triangle::samplePointOnSurface, src/primitives/Triangle.cpp:114-119,159-161)
//   sphere     from outside: a uniform direction in the cone the sphere
//              subtends, pdf 1 / (2 pi (1 - cos theta_max)); from inside: a
//              uniform point on the surface.  The reference returns the
//              nearest surface point with a mis-scaled area pdf
//              (src/primitives/Sphere.cpp:46-67); the cone is the unbiased form.
//   point      a delta light at the point (Point::samplePointOnSurface,
//              src/primitives/Point.cpp:14-26) of intensity = the emitter's
//              radiance, arriving as intensity / d^2 (the reference returns
//              the intensity with pdf 1 and no falloff).
// Returns true for a delta light (no MIS weight).
template <typename R>
NE_DEV R sphere_cone_width(R dc2, R r2) {  // 1 - cos theta_max, cancellation-free
  const R sin2 = r2 / dc2;
  return sin2 / (R(1) + xsqrt(xmax(R(0), R(1) - sin2)));
}
template <typename R>
NE_DEV bool sample_emitter(const DevScene& s, uint32_t code, Vec3<R> x, R ua, R ub, Vec3<R>& wi, R& dist, R& pl,
                           Vec3<R>& Le) {
  const int type = int(code >> 30);
  const uint32_t slot = code & 0x3FFFFFFFu;
  int mat;
  if (type == kLightPoint) {
    const float4 pt = s.points[slot];
    const Vec3<R> dv = v3f4<R>(pt) - x;
    const R d2 = dot(dv, dv);
    dist = xsqrt(d2);
    wi = dv / dist;
    pl = d2 > R(0) ? R(1) : R(0);
    Le = v3f<R>(s.mats + 12 * size_t(__float_as_int(pt.w)) + 6) / d2;
    return true;
  }
  if (type == PRIM_RECT) {
    const float4* rc = s.rects + 4 * size_t(slot);
    const float4 f0 = rc[0], f1 = rc[1], f2 = rc[2], f3 = rc[3];
    const Vec3<R> p0 = v3f4<R>(f0), eu = v3f4<R>(f1), ev = v3f4<R>(f2), ln = v3f4<R>(f3);
    const Vec3<R> y = p0 + ua * eu + ub * ev;
    const Vec3<R> dv = y - x;
    const R d2 = dot(dv, dv);
    dist = xsqrt(d2);
    wi = dv / dist;
    const R cosl = xfabs(dot(ln, wi));
    const R area = length(cross(eu, ev));
    mat = __float_as_int(f0.w);
    pl = cosl > R(0) ? d2 / (cosl * area) : R(0);
  } else if (type == PRIM_SPHERE) {
    const float4 sp = s.sph[slot];
    const Vec3<R> c = v3f4<R>(sp);
    const R r = R(sp.w), r2 = r * r;
    const Vec3<R> oc = c - x;
    const R dc2 = dot(oc, oc);
    mat = s.sph_mat[slot];
    if (dc2 > r2) {
      const R dc = xsqrt(dc2);
      const Vec3<R> w = oc / dc;
      const R wid = sphere_cone_width(dc2, r2);
      const R q = ua * wid;  // 1 - cos theta
      const R ct = R(1) - q, st = xsqrt(xmax(R(0), q * (R(2) - q)));
      const R ph = R(kTWO_PI) * ub;
      Vec3<R> u, v;
      onb(w, u, v);
      R sp, cp;
      xsincos(ph, sp, cp);
      wi = u * (st * cp) + v * (st * sp) + w * ct;
      dist = dc * ct - xsqrt(xmax(R(0), r2 - dc2 * st * st));
      pl = R(1) / (R(kTWO_PI) * wid);
    } else {
      const R z = R(1) - R(2) * ua, rr = xsqrt(xmax(R(0), R(1) - z * z)), ph = R(kTWO_PI) * ub;
      R sp, cp;
      xsincos(ph, sp, cp);
      const Vec3<R> nl = v3(rr * cp, rr * sp, z);
      const Vec3<R> dv = (c + r * nl) - x;
      const R d2 = dot(dv, dv);
      dist = xsqrt(d2);
      wi = dv / dist;
      const R cosl = xfabs(dot(nl, wi));
      pl = cosl > R(0) ? d2 / (cosl * R(2.0 * kTWO_PI) * r2) : R(0);
    }
  } else {
    const float4* tp = s.tris + 3 * size_t(slot);
    const float4 a = tp[0];
    const Vec3<R> v0 = v3f4<R>(a), e1 = v3f4<R>(tp[1]), e2 = v3f4<R>(tp[2]);
    const R su = xsqrt(ua);
    const Vec3<R> y = v0 + (ub * su) * e1 + (su * (R(1) - ub)) * e2;
    const Vec3<R> dv = y - x;
    const R d2 = dot(dv, dv);
    dist = xsqrt(d2);
    wi = dv / dist;
    const Vec3<R> cr = cross(e1, e2);
    const R len = length(cr);
    const R cosl = xfabs(dot(cr, wi)) / len;
    mat = __float_as_int(a.w);
    pl = cosl > R(0) ? d2 / (cosl * (R(0.5) * len)) : R(0);
  }
  Le = v3f<R>(s.mats + 12 * size_t(mat) + 6);
  return false;
}
// Solid-angle pdf with which sample_emitter picks the point that the ray
// (o, d) hits at distance t on emitter (type, slot) with normal n.
template <typename R>
NE_DEV R emitter_hit_pdf(const DevScene& s, int type, uint32_t slot, Vec3<R> o, Vec3<R> d, R t, Vec3<R> n) {
  const R cosl = xfabs(dot(n, d));
  if (type == PRIM_RECT) {
    const R area = R(s.rects[4 * size_t(slot) + 1].w);
    return (t * t) / (cosl * area);
  }
  if (type == PRIM_SPHERE) {
    const float4 sp = s.sph[slot];
    const R r2 = R(sp.w) * R(sp.w);
    const Vec3<R> oc = v3f4<R>(sp) - o;
    const R dc2 = dot(oc, oc);
    if (dc2 > r2) return R(1) / (R(kTWO_PI) * sphere_cone_width(dc2, r2));
    return (t * t) / (cosl * R(2.0 * kTWO_PI) * r2);
  }
  const float4* tp = s.tris + 3 * size_t(slot);
  const R len = length(cross(v3f4<R>(tp[1]), v3f4<R>(tp[2])));
  return (t * t) / (cosl * (R(0.5) * len));
}

// Shadow transmittance: surface occlusion, then ratio tracking through the
// volume over [0, tmax].
// SURF = false: a scene without surfaces (no BVH): only the medium attenuates.
template <typename R, bool C, bool SURF>
NE_DEV R transmittance_t(const DevScene& s, Vec3<R> o, Vec3<R> d, R tmax, PathRng& rng, uint32_t* lds, int stride,
                         Count<C>& cnt) {
  cnt.shadow();
  if constexpr (SURF) {
    SurfHit<R> h;
    if (traverse<R, true, C>(s, o, d, R(1e-4), tmax, lds, stride, h, cnt)) return R(0);
  }
  if (!s.has_volume) return R(1);
  Vec3<R> oo = xf_point<R>(s.inv, o), dd = xf_vec<R>(s.inv, d);
  R ta, tb;
  if (!box_interval<R>(s, oo, dd, ta, tb)) return R(1);
  ta = xmax(ta, R(0));
  tb = xmin(tb, tmax);
  if (ta >= tb) return R(1);
  return ratio_track<R, C>(s, o, d, ta, tb, rng, cnt);
}

template <typename R, bool C>
NE_DEV R transmittance(const DevScene& s, Vec3<R> o, Vec3<R> d, R tmax, PathRng& rng, uint32_t* lds, int stride,
                       Count<C>& cnt) {
  return transmittance_t<R, C, true>(s, o, d, tmax, rng, lds, stride, cnt);
}

// ---------------------------------------------------------------------------
// One path (corrected semantics).  The event and random-draw order is the
// oracle's Li (camera jitter is drawn by the caller).
// ---------------------------------------------------------------------------
template <typename R>
struct PathState {
  Vec3<R> o, d;     // current ray
  Vec3<R> L, beta;  // radiance gathered, throughput
  R prev_pdf;       // pdf of the direction that produced the ray (emission MIS)
  bool prev_delta;  // the ray left the camera (no MIS on what it hits)
  int depth;        // scatter vertices so far
};

template <typename R>
NE_DEV void path_init(PathState<R>& p, Vec3<R> o, Vec3<R> d) {
  p.o = o;
  p.d = d;
  p.L = v3s(R(0));
  p.beta = v3s(R(1));
  p.prev_pdf = R(0);
  p.prev_delta = true;
  p.depth = 0;
}

// One bounce of the path: the segment from p.o along p.d (surface hit and/or
// medium collision), emission / escape, and at a scatter vertex next-event
// estimation plus the sampled continuation.  Returns true when the path has
// ended (p.L is final).  trace_path loops over it; the persistent batch
// kernel interleaves bounces of different paths in one wave.  SURF = false
// compiles the surface code out, for scenes with neither surfaces nor
// emitters (a volume lit by the sun and the environment): the same events,
// arithmetic and draws, with fewer registers.
// GUIDED (FP32): medium vertices at depth < gu.max_depth sample the one-sample
// MIS mixture (1 - prob) HG + prob proposal, the proposal read from the
// guide's tabulation (gu.cache; nullptr: plain HG), with the wavefront's draw
// order (selector, two for the direction, one for the lobe), and the light
// sample's MIS weight uses the mixture pdf.
template <typename R, bool C, bool SURF = true, bool GUIDED = false>
NE_DEV bool path_bounce(const DevScene& s, PathState<R>& p, int max_bounces, PathRng& rng, uint32_t* lds,
                        int stride, Count<C>& cnt, const GuideUse* gu = nullptr) {
  Vec3<R> o = p.o, d = p.d, L = p.L, beta = p.beta;
  const R prev_pdf = p.prev_pdf;
  const bool prev_delta = p.prev_delta;
  const int depth = p.depth;
  const R tmin = R(1e-4);
  const R g = R(s.g);
  const R nl = R(s.n_lights);
  {
    cnt.ext();
    SurfHit<R> hit;
    bool hs = false;
    if constexpr (SURF) hs = traverse<R, false, C>(s, o, d, tmin, xinf<R>(), lds, stride, hit, cnt);
    R tS = hs ? hit.t : xinf<R>();
    bool medium = false;
    Vec3<R> x, n;
    int mat = 0;
    if (s.has_volume) {
      Vec3<R> oo = xf_point<R>(s.inv, o), dd = xf_vec<R>(s.inv, d);
      R ta, tb;
      if (box_interval<R>(s, oo, dd, ta, tb)) {
        ta = xmax(ta, R(0));
        tb = xmin(tb, tS);
        R th;
        if (ta < tb && delta_track<R, C>(s, o, d, ta, tb, rng, cnt, th)) { medium = true; x = o + th * d; }
      }
    }
    if (!medium) {
      if (!hs) {
        if (s.env_type != 0) {
          R w = prev_delta ? R(1) : power_heuristic(prev_pdf, env_pdf<R>(s, d));
          L = L + beta * env_le<R>(s, d) * w;
        }
        p.L = L;
        return true;
      }
      if constexpr (SURF) {
        hit_normal<R>(s, hit, o, d, n, mat);
        const float* m = s.mats + 12 * size_t(mat);
        if (m[0] == 1.0f) {  // emitter
          Vec3<R> Le = v3f<R>(m + 6);
          R w = R(1);
          if (!prev_delta) w = power_heuristic(prev_pdf, emitter_hit_pdf<R>(s, hit.type, hit.slot, o, d, hit.t, n));
          L = L + beta * Le * w;
          p.L = L;
          return true;
        }
        x = o + hit.t * d;
      }
    }
    const bool med = !SURF || medium;  // without surfaces every vertex is a medium collision
    if (depth >= max_bounces) {
      p.L = L;
      return true;
    }
    // ---- vertex: next-event estimation
    Vec3<R> wo = -d;
    Glossy<R> gm;
    Vec3<R> ns, ss, ts;
    if (med) {
      beta = beta * v3(R(s.sig_s[0]) / R(s.sig_t[0]), R(s.sig_s[1]) / R(s.sig_t[1]), R(s.sig_s[2]) / R(s.sig_t[2]));
    } else {
      const float* m = s.mats + 12 * size_t(mat);
      gm.albedo = v3f<R>(m + 1);
      R rough = R(m[4]);
      gm.alpha = rough * rough;
      gm.metallic = R(m[5]);
      ns = dot(n, wo) < R(0) ? -n : n;
      onb(ns, ss, ts);
    }
    // the guided proposal of this vertex (tabulated entry), if any
    const float* mix = nullptr;
    if constexpr (GUIDED) {
      if (med && gu->cache.mix && depth < gu->max_depth)
        mix = guide_cache_at(gu->cache, xf_point<float>(s.inv, x) + v3s(0.5f), d);
    }
    R ul = R(rng.next()), ua = R(rng.next()), ub = R(rng.next());
    if (s.n_lights > 0) {
      int li = int(ul * nl);
      li = li < s.n_lights - 1 ? li : s.n_lights - 1;
      Vec3<R> wi, Le = v3s(R(0));
      R pl = R(0), dist = xinf<R>();
      bool delta = false;
      const int env_slot = s.has_sun ? 1 : 0;
      if (s.has_sun && li == 0) {
        wi = -v3f<R>(s.sun_dir);
        Le = v3f<R>(s.sun_E);
        pl = R(1);
        delta = true;
      } else if (s.env_type != 0 && li == env_slot) {
        env_sample<R>(s, ua, ub, wi, pl, Le);
      } else if constexpr (SURF) {
        const uint32_t code = s.emitters[li - env_slot - (s.env_type != 0 ? 1 : 0)];
        delta = sample_emitter<R>(s, code, x, ua, ub, wi, dist, pl, Le);
      }
      if (pl > R(0)) {
        Vec3<R> f;
        R pf;
        if (med) {
          R p = hg_p(g, dot(d, wi));
          f = v3s(p);
          pf = p;
        } else {
          Vec3<R> wol = to_lcs(wo, ns, ss, ts), wil = to_lcs(wi, ns, ss, ts);
          gm.eval_pdf(wol, wil, f, pf);
          f = f * xfabs(wil.z);
        }
        if (!is_black(f)) {
          R tmax = xisinf(dist) ? xinf<R>() : dist * R(1 - 1e-4);
          PathRng srng = rng.shadow_stream(depth);
          R tr = transmittance_t<R, C, SURF>(s, x, wi, tmax, srng, lds, stride, cnt);
          if (tr > R(0)) {
            R pm = pf;
            if constexpr (GUIDED) {
              if (mix) pm = (1.0f - gu->prob) * pf + gu->prob * vmf_mixture_pdf(mix, wi);
            }
            R w = delta ? R(1) : power_heuristic(pl, pm);
            L = L + beta * f * Le * (tr * w * nl / pl);
          }
        }
      }
    }
    // ---- continue the path
    if (GUIDED && mix) {
      if constexpr (GUIDED) {
        const float us = rng.next(), u0 = rng.next(), u1 = rng.next(), u2 = rng.next();
        const Vec3<float> wi = us < gu->prob ? vmf_mixture_sample(mix, u2, u0, u1) : hg_sample(g, d, u0, u1);
        const float ph = hg_p(g, dot(d, wi));
        const float pm = (1.0f - gu->prob) * ph + gu->prob * vmf_mixture_pdf(mix, wi);
        if (!(pm > 0.0f)) {
          p.L = L;
          return true;
        }
        p.prev_pdf = pm;
        p.d = wi;
        p.beta = beta * (ph / pm);
      }
    } else if (med) {
      R u0 = R(rng.next()), u1 = R(rng.next());
      Vec3<R> wi = hg_sample(g, d, u0, u1);
      p.prev_pdf = hg_p(g, dot(d, wi));
      p.d = wi;
      p.beta = beta;  // scattering albedo applied at the vertex
    } else {
      R u0 = R(rng.next()), u1 = R(rng.next()), u2 = R(rng.next());
      Vec3<R> wol = to_lcs(wo, ns, ss, ts);
      Vec3<R> wil = gm.sample(wol, u0, u1, u2);
      Vec3<R> fe;
      R pf;
      gm.eval_pdf(wol, wil, fe, pf);
      if (!(pf > R(0))) {
        p.L = L;
        return true;
      }
      Vec3<R> f = fe * xfabs(wil.z);
      if (is_black(f)) {
        p.L = L;
        return true;
      }
      p.beta = beta * f / pf;
      p.prev_pdf = pf;
      p.d = to_world(wil, ns, ss, ts);
    }
    p.L = L;
    p.prev_delta = false;
    p.o = x;
    p.depth = depth + 1;
  }
  return false;
}

// One whole path (the oracle's Li event and draw order; camera jitter is drawn
// by the caller).
template <typename R, bool C, bool SURF = true, bool GUIDED = false>
NE_DEV Vec3<R> trace_path(const DevScene& s, Vec3<R> o, Vec3<R> d, int max_bounces, PathRng& rng, uint32_t* lds,
                          int stride, Count<C>& cnt, const GuideUse* gu = nullptr) {
  PathState<R> p;
  path_init(p, o, d);
  while (!path_bounce<R, C, SURF, GUIDED>(s, p, max_bounces, rng, lds, stride, cnt, gu)) {
  }
  return p.L;
}

// Order-independent film accumulation: 64-bit fixed point, 2^-30 radiance
// units per count; integer atomics make the sum independent of the order in
// which paths finish.
constexpr float kFixedScale = 1073741824.0f;
NE_DEV void film_add(long long* film, uint32_t pix, float r, float g, float b) {
  long long* f = film + size_t(pix) * 3;
  if (r != 0.0f) atomicAdd((unsigned long long*)&f[0], (unsigned long long)__float2ll_rn(r * kFixedScale));
  if (g != 0.0f) atomicAdd((unsigned long long*)&f[1], (unsigned long long)__float2ll_rn(g * kFixedScale));
  if (b != 0.0f) atomicAdd((unsigned long long*)&f[2], (unsigned long long)__float2ll_rn(b * kFixedScale));
}

}  // namespace ne
