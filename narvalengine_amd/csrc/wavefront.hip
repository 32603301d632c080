// wavefront.hip -- wavefront path tracer for gfx950 (FP32 product path).
//
// Path state lives in HBM as two structure-of-arrays buffers A and B of P
// slots, split into chunks of 256 slots.  Every stage kernel runs one
// 256-thread block per chunk, thread t owning slot t of its chunk, and the
// live paths of a chunk are kept packed at its front, so all state accesses
// are coalesced.  One iteration:
//
//   plan      one block scans the chunks' free slots and hands each chunk a
//             contiguous range of new path ids (no atomics: the assignment is
//             a deterministic prefix sum)
//   regen     fills the free tail of each chunk with camera rays for its ids
//             (ids -> pixels in 8x8 tiles: neighbouring slots trace
//             neighbouring pixels)
//   trace     BVH closest hit, LDS short stack            (scenes with surfaces)
//   medium    delta tracking through the majorant grid    (scenes with a volume)
//   shade     emission / escape / depth limit, then the phase-function or
//             glossy vertex: NEE light sample -> shadow queue of the chunk,
//             continuation -> chunk of B (block-local compaction in LDS),
//             ended paths -> film
//   shadow    any-hit occlusion and/or ratio tracking; the NEE contribution
//             goes to the continuing path in B or to the film
//   swap A/B
//
// Same-address global atomics serialise across the XCDs (about 11 ns each on
// MI355X), so nothing here appends to a shared queue: compaction is a
// per-block prefix sum and the only global atomics are film additions, spread
// over pixels.  The film is accumulated in 64-bit fixed point (2^-30 units):
// the sum does not depend on completion order.  Random numbers follow the
// megakernel's per-path Philox stream (each path carries the number of draws
// consumed) and shadow rays use their vertex's sub-stream, so both modes trace
// the same paths with the same arithmetic.
#include <algorithm>
#include <cstring>

#include "ne_guide.h"
#include "ne_launch.h"
#include "ne_trace.h"

namespace ne {

constexpr uint32_t kToFilm = 0x80000000u;      // shadow target flag: add to film pixel
constexpr int kPlanThreads = 1024;

// Position of this thread's item among the block's selected items (in thread
// order) and the block total.  Every thread of the block must call it.
__device__ __forceinline__ uint32_t block_compact(bool pred, uint32_t* wsum, uint32_t& total) {
  const unsigned long long mask = __ballot(pred);
  const int lane = __lane_id(), wave = int(threadIdx.x >> 6);
  if (lane == 0) wsum[wave] = uint32_t(__popcll(mask));
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (int k = 0; k < int(kChunk / 64); ++k) {
    before += k < wave ? wsum[k] : 0u;
    all += wsum[k];
  }
  __syncthreads();  // wsum is reused by the next call
  total = all;
  return before + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
}

// Spatial order of a chunk's rays: thread t of the block gets the chunk slot
// of the t-th ray in Morton order of its origin (8 bits per axis of the
// volume's unit cube, clamped), live rays (t < n) first.  The 23-bit key and
// the slot share one word, sorted bitonically in LDS.  Every thread of the
// block must call it.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // 8 bits -> every third bit
  v = (v | (v << 8)) & 0x0F00Fu;
  v = (v | (v << 4)) & 0x0C30C3u;
  v = (v | (v << 2)) & 0x249249u;
  return v;
}
__device__ __forceinline__ uint32_t chunk_spatial_slot(const DevScene& s, bool live, Vec3<float> o, uint32_t* skey) {
  const uint32_t t = threadIdx.x;
  uint32_t key = 0xFFFFFFu;  // free slots sort last
  if (live) {
    const Vec3<float> u = xf_point<float>(s.inv, o) + v3s(0.5f);
    const uint32_t x = uint32_t(xmin(xmax(u.x, 0.0f), 0.999f) * 256.0f);
    const uint32_t y = uint32_t(xmin(xmax(u.y, 0.0f), 0.999f) * 256.0f);
    const uint32_t z = uint32_t(xmin(xmax(u.z, 0.0f), 0.999f) * 256.0f);
    key = (spread3(x) | (spread3(y) << 1) | (spread3(z) << 2)) >> 1;  // 23 bits, below the free marker
  }
  skey[t] = (key << 8) | t;
  __syncthreads();
  for (uint32_t k = 2; k <= kChunk; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      const uint32_t p = t ^ j;
      if (p > t) {
        const uint32_t x0 = skey[t], x1 = skey[p];
        if ((x0 > x1) == ((t & k) == 0)) { skey[t] = x1; skey[p] = x0; }
      }
      __syncthreads();
    }
  return skey[t] & 0xFFu;
}

// Sums a per-thread counter over the block and adds it once to *dst.
template <bool C>
__device__ __forceinline__ void block_count(unsigned long long* dst, uint32_t v, uint32_t* red) {
  if constexpr (C) {
    unsigned long long s = v;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
    if (__lane_id() == 0) red[threadIdx.x >> 6] = uint32_t(s);
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (int k = 0; k < int(kChunk / 64); ++k) t += red[k];
      if (t) atomicAdd(dst, t);
    }
    __syncthreads();
  }
}

__global__ void k_wf_init(WfState w, uint32_t nchunks) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nchunks) w.cnt_paths[0][c] = 0;
  if (c < (nchunks + kPlanThreads - 1) / kPlanThreads) w.plan_lb[c] = 0ull;  // epoch 0: no word yet
  if (c == 0) {
    *w.next_path = 0;
    w.status[1] = 0u;
    w.status[2] = 0u;
    w.status[3] = 0u;
  }
}

// id / n and id % n for a path id and 0 < n < 2^32 with a quotient below
// 2^32.  The correctly rounded double quotient is within 2^-20 of the true
// one, so its floor is the true quotient or one off, and one correction step
// each way makes it exact (a few instructions instead of a 64-bit integer
// division).
__device__ __forceinline__ void divmod_id(unsigned long long id, uint32_t n, uint32_t& q, uint32_t& r) {
  const double e = floor(double(id) / double(n));
  long long qq = (long long)e;
  long long rr = (long long)id - qq * (long long)n;
  if (rr < 0) { qq -= 1; rr += n; }
  if (rr >= (long long)n) { qq += 1; rr -= n; }
  q = uint32_t(qq);
  r = uint32_t(rr);
}

// path id -> (pixel, sample): one sample index at a time over the rect; full
// 8x8 tiles first (row-major tile order), then the right strip, then the
// bottom strip (row-major): a bijection onto the rect.  (px, py): the pixel's
// coordinates within the rect.  Rects hold at most 2^30 pixels and renders
// fewer than 2^32 samples per pixel.
__device__ __forceinline__ void path_pixel(const WfArgs& a, unsigned long long id, uint32_t& pix, uint32_t& smp,
                                           uint32_t& px, uint32_t& py) {
  const uint32_t rw = uint32_t(a.x1 - a.x0), rh = uint32_t(a.y1 - a.y0);
  uint32_t p;
  divmod_id(id, rw * rh, smp, p);
  const uint32_t full_rows = rh / 8, full_cols = rw / 8;
  const uint32_t tiles_full = full_rows * full_cols;
  if (p < tiles_full * 64) {
    const uint32_t t = p / 64, l = p % 64;
    const uint32_t ty = t / full_cols;
    px = (t - ty * full_cols) * 8 + (l & 7);
    py = ty * 8 + (l >> 3);
  } else {
    uint32_t r = p - tiles_full * 64;
    const uint32_t strip_w = rw - full_cols * 8;
    const uint32_t right = strip_w * full_rows * 8;
    if (r < right) {
      px = full_cols * 8 + r % strip_w;
      py = r / strip_w;
    } else {
      r -= right;
      px = r % rw;
      py = full_rows * 8 + r / rw;
    }
  }
  pix = uint32_t(a.y0 + int(py)) * uint32_t(a.W) + uint32_t(a.x0 + int(px));
}

// Per-sample radiance: with a.sbuf, a path's radiance goes to its own slot
// (sample - sample_offset) * rect pixels + rect pixel, one plain store when
// the path ends; k_wf_flush_samples then sums each pixel's samples in sample
// order in FP64 and adds the fixed-point film, which still takes the final
// vertices' light samples.  Without a.sbuf the fixed-point film takes
// everything.  Both sums are independent of completion order.
__device__ __forceinline__ size_t sample_slot(const WfArgs& a, uint32_t pix, uint32_t smp) {
  const uint32_t px = pix % uint32_t(a.W), py = pix / uint32_t(a.W);
  const size_t rw = size_t(a.x1 - a.x0), rh = size_t(a.y1 - a.y0);
  return (size_t(smp - a.sample_offset) * rh + size_t(py - uint32_t(a.y0))) * rw + size_t(px - uint32_t(a.x0));
}
__device__ __forceinline__ void sample_store(const WfArgs& a, uint32_t pix, uint32_t smp, float r, float g,
                                             float b) {
  float* q = a.sbuf + 3 * sample_slot(a, pix, smp);
  q[0] = r; q[1] = g; q[2] = b;
}

// The plan: an exclusive scan of the chunks' free slots gives each chunk
// its first new path id and count; it advances the work counter and writes
// the number of paths in flight this iteration to status[0].  One block per
// kPlanThreads chunks (a thread per chunk, loads and stores coalesced): the
// block scans its own chunks (wave shuffles, then the wave totals in LDS)
// and takes the free slots of all earlier blocks by decoupled look-back.
// Each block publishes one 64-bit word, (epoch << 32) | (kind << 31) |
// count: first its own total (kind 0), then the total of every chunk up to
// its last (kind 1).  Wave 0 reads 64 predecessors' words per round, nearest
// first, and adds counts up to the nearest inclusive total.  Words of
// another launch carry another epoch and are waited for; every block's
// predecessors run before or beside it, so the wait ends, and a bounded
// spin turns a word that never comes into status[3] instead of a hang.
// a.plan_stall (a test of that path) makes the readers wait for the next
// epoch, whose words never come, with a short spin.
constexpr uint32_t kPlanSpinLimit = 1u << 22, kPlanStallSpins = 1u << 10;
__device__ __forceinline__ unsigned long long plan_word(const unsigned long long* lb, int j, uint32_t epoch,
                                                        uint32_t* fail, uint32_t limit) {
  if (j < 0) return (1ull << 31);  // before the first block: an inclusive total of 0
  unsigned long long v = __hip_atomic_load(lb + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (uint32_t k = 0; uint32_t(v >> 32) != epoch; ++k) {
    if (k == limit) {
      __hip_atomic_store(fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return (1ull << 31);
    }
    __builtin_amdgcn_s_sleep(1);
    v = __hip_atomic_load(lb + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return v;
}
constexpr int kPlanWaves = kPlanThreads / 64;
__global__ __launch_bounds__(kPlanThreads) void k_wf_plan(WfState w, WfArgs a) {
  __shared__ uint32_t wtot[kPlanWaves];  // wave totals, then their exclusive prefix
  __shared__ uint32_t btot;
  __shared__ unsigned long long bexcl;  // free slots of the chunks before this block
  __shared__ unsigned long long sbase;  // the work counter, read once before this block publishes
  const uint32_t NC = a.nchunks, b = blockIdx.x;
  const uint32_t c = b * uint32_t(kPlanThreads) + threadIdx.x;
  const int lane = __lane_id(), wave = int(threadIdx.x >> 6);
  // every block reads the counter before its first word, and the last block
  // advances it only after it has a word from every block before it
  if (threadIdx.x == 0) sbase = *w.next_path;
  const uint32_t fr = c < NC ? kChunk - w.cnt_paths[a.cur][c] : 0u;
  uint32_t incl = fr;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  if (lane == 63) wtot[wave] = incl;
  __syncthreads();
  if (wave == 0) {
    const uint32_t x = lane < kPlanWaves ? wtot[lane] : 0u;
    uint32_t sc = x;
#pragma unroll
    for (int o = 1; o < kPlanWaves; o <<= 1) {
      const uint32_t v = __shfl_up(sc, o);
      if (lane >= o) sc += v;
    }
    if (lane < kPlanWaves) wtot[lane] = sc - x;
    if (lane == kPlanWaves - 1) btot = sc;
  }
  __syncthreads();
  if (wave == 0) {
    const uint32_t agg = btot, ep = a.plan_epoch;
    const uint32_t want = ep + a.plan_stall, limit = a.plan_stall ? kPlanStallSpins : kPlanSpinLimit;
    unsigned long long* lb = w.plan_lb;
    if (lane == 0)
      __hip_atomic_store(lb + b, ((unsigned long long)ep << 32) | (b == 0 ? (1ull << 31) : 0ull) | agg,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long excl = 0;
    for (int j = int(b) - 1; j >= 0; j -= 64) {
      const unsigned long long v = plan_word(lb, j - lane, want, w.status + 3, limit);
      const unsigned long long pm = __ballot(((v >> 31) & 1ull) != 0ull);
      const int f = pm ? __ffsll((long long)pm) - 1 : 63;  // nearest inclusive total
      unsigned long long cnt = lane <= f ? (v & 0x7FFFFFFFull) : 0ull;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
      excl += cnt;
      if (pm) break;
    }
    if (lane == 0) {
      if (b > 0)
        __hip_atomic_store(lb + b, ((unsigned long long)ep << 32) | (1ull << 31) | (excl + agg),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      bexcl = excl;
    }
  }
  __syncthreads();
  const unsigned long long base = sbase;
  if (c < NC) {
    const unsigned long long off = base + bexcl + wtot[wave] + incl - fr;
    const unsigned long long left = off < a.total_paths ? a.total_paths - off : 0ull;
    w.regen_base[c] = off;
    w.regen_m[c] = uint32_t(left < fr ? left : fr);
  }
  if (b == gridDim.x - 1 && threadIdx.x == 0) {  // the last block holds the total
    const unsigned long long total_free = bexcl + btot;
    const unsigned long long left = base < a.total_paths ? a.total_paths - base : 0ull;
    const unsigned long long m = left < total_free ? left : total_free;
    *w.next_path = base + m;
    const unsigned long long inflight = (unsigned long long)NC * kChunk - total_free + m;
    w.status[0] = uint32_t(inflight < 0xFFFFFFFFull ? inflight : 0xFFFFFFFFull);
  }
}

// New paths fill the free tail of chunk c: thread t < m writes the camera ray
// of path regen_base[c] + t into slot n0 + t.
// Path id -> its camera ray, written as slot's record; also returned (ray
// origin and direction, pixel, sample, draws consumed).
__device__ __forceinline__ void regen_path(const WfArgs& a, const DevCamera& cam, const WfPaths& A, uint32_t slot,
                                           unsigned long long id, Vec3<float>& o, Vec3<float>& d, uint32_t& pix,
                                           uint32_t& smp, uint32_t& rngn) {
  uint32_t rx, ry;
  path_pixel(a, id, pix, smp, rx, ry);
  smp += a.sample_offset;
  PathRng rng;
  rng.init(a.seed, pix, smp, a.job);
  const float ju = rng.next(), jv = rng.next();
  const uint32_t px = uint32_t(a.x0) + rx, py = uint32_t(a.y0) + ry;  // pix % W, pix / W
  const float u = (float(px) + ju) / float(a.W), v = (float(py) + jv) / float(a.H);
  const Vec3<float> pos = v3f<float>(cam.pos), ll = v3f<float>(cam.lower_left);
  const Vec3<float> hz = v3f<float>(cam.horizontal), vt = v3f<float>(cam.vertical);
  const Vec3<float> dir = -normalize(ll + u * hz + v * vt - pos);
  rngn = rng.consumed();
  A.ox[slot] = pos.x; A.oy[slot] = pos.y; A.oz[slot] = pos.z;
  A.dx[slot] = dir.x; A.dy[slot] = dir.y; A.dz[slot] = dir.z;
  A.bx[slot] = 1.0f; A.by[slot] = 1.0f; A.bz[slot] = 1.0f;
  A.Lx[slot] = 0.0f; A.Ly[slot] = 0.0f; A.Lz[slot] = 0.0f;
  A.pix[slot] = pix; A.smp[slot] = smp;
  A.rngn[slot] = rngn;
  A.depth[slot] = kPrevDelta;  // depth 0, camera ray (delta)
  A.prev_pdf[slot] = 0.0f;
  if (a.gstate[a.cur]) a.gstate[a.cur][slot] = 0u;
  o = pos;
  d = dir;
}
__device__ __forceinline__ void regen_chunk(const WfState& w, const WfArgs& a, const DevCamera& cam, uint32_t c,
                                            uint32_t t, uint32_t n0, uint32_t m) {
  if (t < m) {
    Vec3<float> o, d;
    uint32_t pix, smp, rngn;
    regen_path(a, cam, w.paths[a.cur], c * kChunk + n0 + t, w.regen_base[c] + t, o, d, pix, smp, rngn);
  }
}

// Paths of chunk c in the current buffer: the live count, plus the new paths
// when the medium stage made them this iteration (it leaves the count).
__device__ __forceinline__ uint32_t chunk_paths(const WfState& w, const WfArgs& a, uint32_t c) {
  return w.cnt_paths[a.cur][c] + (a.regen_fused ? w.regen_m[c] : 0u);
}

__global__ __launch_bounds__(kChunk) void k_wf_regen(WfState w, WfArgs a, DevCamera cam) {
  const uint32_t c = blockIdx.x, t = threadIdx.x;
  const uint32_t n0 = w.cnt_paths[a.cur][c], m = w.regen_m[c];
  regen_chunk(w, a, cam, c, t, n0, m);
  __syncthreads();  // every thread has read n0
  if (t == 0) w.cnt_paths[a.cur][c] = n0 + m;
}

// ---- lane refill -------------------------------------------------------------
// Traversal lengths are heavy-tailed (rays grazing dense geometry visit many
// nodes), so a wave that holds one ray per lane waits for its slowest lane.
// The BVH kernels below give each 256-thread block a pool of a.refill chunks
// (sized on the host so the grid is one round of resident blocks) and run one
// flattened loop: every iteration, lanes without a ray take the next one from
// the pool (one LDS atomic per wave) and every lane with a ray does one node
// visit.  A lane leaves the loop only when the pool is empty.

// Per-block pool over the chunks blockIdx.x + j * gridDim.x (j < n): chunks
// map to image regions, so interleaving spreads expensive regions over all
// blocks.  Items are numbered chunk by chunk, pre[j] = first item of chunk j,
// pre[n] = total.
struct RefillPool {
  uint32_t pre[kMaxRefill + 1];
  uint32_t n, next;
};

__device__ __forceinline__ uint32_t pool_chunk(uint32_t j) { return blockIdx.x + j * gridDim.x; }

__device__ __forceinline__ void pool_init(RefillPool& p, const uint32_t* cnt, uint32_t nchunks) {
  if (threadIdx.x == 0) {
    uint32_t n = 0, acc = 0;
    while (n < kMaxRefill && pool_chunk(n) < nchunks) {
      p.pre[n] = acc;
      acc += cnt[pool_chunk(n)];
      ++n;
    }
    p.pre[n] = acc;
    p.n = n;
    p.next = 0;
  }
  __syncthreads();
}

// Lanes with need == true take consecutive items (wave-aggregated LDS atomic).
// Returns the slot (chunk * kChunk + position) or kNoHit when the pool is empty.
// Every lane still in the caller's loop must call it.
__device__ __forceinline__ uint32_t pool_fetch(RefillPool& p, bool need) {
  const unsigned long long m = __ballot(need);
  if (m == 0) return kNoHit;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(&p.next, uint32_t(__popcll(m)));
  base = __shfl(base, leader);
  if (!need) return kNoHit;
  const uint32_t k = base + uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
  const uint32_t n = p.n;
  if (k >= p.pre[n]) return kNoHit;
  // largest j < n with pre[j] <= k (pre is non-decreasing, pre[0] = 0)
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (p.pre[mid] <= k) lo = mid; else hi = mid - 1;
  }
  return pool_chunk(lo) * kChunk + (k - p.pre[lo]);
}

// Refill vote of the lanes still in a refill loop: true (on every such lane)
// once at least `min_idle` of them, or all of them, have no ray.  Idle lanes
// wait in between, so the refill path runs on fewer steps.
__device__ __forceinline__ bool refill_due(bool have, int min_idle) {
  const unsigned long long idle = __ballot(!have), looping = __ballot(true);
  return idle != 0ull && (__popcll(idle) >= min_idle || idle == looping);
}

// Closest surface hit of every live ray: pure BVH traversal (no random
// numbers), lane refill over the block's pool.
template <bool C>
__global__ __launch_bounds__(kChunk) void k_wf_trace(WfState w, WfArgs a, DevScene s) {
  __shared__ uint32_t lds_stack[kShortStack * kChunk];
  __shared__ uint32_t red[kChunk / 64];
  __shared__ RefillPool pool;
  const WfPaths A = w.paths[a.cur];
  pool_init(pool, w.cnt_paths[a.cur], a.nchunks);
  uint32_t* stk = lds_stack + threadIdx.x;
  uint32_t spill[kSpillStack];
  Count<C> cnt;
  TravState<float> st;
  uint32_t slot = kNoHit;
  bool have = false;
  while (true) {
    if (refill_due(have, a.refill_min)) {
      const uint32_t got = pool_fetch(pool, !have);  // every lane still looping calls it
      if (!have) {
        if (got == kNoHit) break;
        slot = got;
        have = true;
        cnt.ext();
        trav_init(st, v3(A.ox[slot], A.oy[slot], A.oz[slot]), v3(A.dx[slot], A.dy[slot], A.dz[slot]), 1e-4f,
                  xinf<float>());
      }
    }
    if (have && trav_step_wave<float, false, C>(s, st, stk, kChunk, spill, cnt)) {
      A.hit_t[slot] = st.found ? st.best : xinf<float>();
      A.hit_ref[slot] = st.found ? ((uint32_t(st.btype) << 30) | st.bslot) : kNoHit;
      A.kind[slot] = st.found ? uint32_t(WF_KIND_SURFACE) : uint32_t(WF_KIND_MISS);
      have = false;
    }
  }
  if constexpr (C) {
    block_count<C>(&a.counters->ext_rays, cnt.n_ext, red);
    block_count<C>(&a.counters->bvh_nodes, cnt.n_node, red);
    block_count<C>(&a.counters->prim_tests, cnt.n_prim, red);
  }
}

// Delta tracking through the volume up to the surface hit (or infinity).
// HAS_SURF: the trace stage ran and set kind / hit_t.
// SPATIAL: the chunk's rays are taken in Morton order of their origins, so a
// wave tracks through neighbouring voxels (each path keeps its own stream:
// same trajectories).
// REGEN (volume-only scenes): the block first makes its chunk's new paths, as
// k_wf_regen would, which saves that stage's launch every iteration.
// SMALL: the scene's grid is under 4 GiB (density lookups by 32-bit offsets).
template <bool C, bool HAS_SURF, bool SPATIAL, bool REGEN, bool SMALL>
__global__ __launch_bounds__(kChunk) void k_wf_medium(WfState w, WfArgs a, DevScene s, DevCamera cam) {
  __shared__ uint32_t red[kChunk / 64];
  __shared__ uint32_t skey[SPATIAL ? kChunk : 1];
  const WfPaths A = w.paths[a.cur];
  const uint32_t c = blockIdx.x, t = threadIdx.x;
  uint32_t n;
  // a new path (REGEN) is made by the thread that tracks it, so its ray stays
  // in registers: no block barrier, no reload; the live count is left as it
  // was (a.regen_fused: later stages add regen_m)
  bool fresh = false;
  Vec3<float> fo = v3s(0.0f), fd = v3s(0.0f);
  uint32_t fpix = 0, fsmp = 0, frngn = 0;
  if constexpr (REGEN) {
    const uint32_t n0 = w.cnt_paths[a.cur][c];
    n = n0 + w.regen_m[c];
    if (t >= n0 && t < n) {
      fresh = true;
      regen_path(a, cam, A, c * kChunk + t, w.regen_base[c] + (t - n0), fo, fd, fpix, fsmp, frngn);
    }
  } else {
    n = w.cnt_paths[a.cur][c];
  }
  uint32_t slot = t;
  if constexpr (SPATIAL) {
    const uint32_t q = c * kChunk + t;
    slot = chunk_spatial_slot(s, t < n, t < n ? v3(A.ox[q], A.oy[q], A.oz[q]) : v3s(0.0f), skey);
  }
  const uint32_t i = c * kChunk + slot;
  Count<C> cnt;
  if (t < n) {
    Vec3<float> o = fo, d = fd;
    float tS = xinf<float>();
    uint32_t ppix = fpix, psmp = fsmp, prngn = frngn;
    if (!fresh) {
      o = v3(A.ox[i], A.oy[i], A.oz[i]);
      d = v3(A.dx[i], A.dy[i], A.dz[i]);
      if (HAS_SURF) tS = A.hit_t[i];
      // the stream's key goes out with the ray, not after the box test (one
      // round trip); the empty asm takes the values here so that the loads
      // are not sunk past the test
      ppix = A.pix[i];
      psmp = A.smp[i];
      prngn = A.rngn[i];
      asm volatile("" ::"v"(o.x), "v"(o.y), "v"(o.z), "v"(d.x), "v"(d.y), "v"(d.z), "v"(tS), "v"(ppix),
                   "v"(psmp), "v"(prngn));
    }
    if (!HAS_SURF) cnt.ext();
    bool medium = false;
    float th = 0.0f;
    const Vec3<float> oo = xf_point<float>(s.inv, o), dd = xf_vec<float>(s.inv, d);
    float ta, tb;
    if (box_interval<float>(s, oo, dd, ta, tb)) {
      ta = xmax(ta, 0.0f);
      tb = xmin(tb, tS);
      if (ta < tb) {
        PathRng rng;
        rng.init(a.seed, ppix, psmp, a.job);
        rng.seek(prngn);
        medium = delta_track<float, C, SMALL>(s, o, d, ta, tb, rng, cnt, th);
        A.rngn[i] = rng.consumed();
      }
    }
    if (medium) {
      A.hit_t[i] = th;
      A.kind[i] = WF_KIND_MEDIUM;
    } else if (!HAS_SURF) {
      A.kind[i] = WF_KIND_MISS;
    }
  }
  if constexpr (C) {
    if (!HAS_SURF) block_count<C>(&a.counters->ext_rays, cnt.n_ext, red);
    block_count<C>(&a.counters->density_lookups, cnt.n_lookup, red);
    block_count<C>(&a.counters->majorant_cells, cnt.n_cell, red);
  }
}

// Scenes with neither surfaces nor a volume: every ray escapes.
__global__ __launch_bounds__(kChunk) void k_wf_all_miss(WfState w, WfArgs a) {
  const uint32_t c = blockIdx.x, t = threadIdx.x;
  if (t < w.cnt_paths[a.cur][c]) w.paths[a.cur].kind[c * kChunk + t] = WF_KIND_MISS;
}

// NEE light sample at x (trace_path's light-selection block).  Returns false
// when no shadow ray is needed.
template <bool MEDIUM>
__device__ __forceinline__ bool nee_sample(const DevScene& s, Vec3<float> x, Vec3<float> d, const Glossy<float>& gm,
                                           Vec3<float> wo, Vec3<float> ns, Vec3<float> ss, Vec3<float> ts, float ul,
                                           float ua, float ub, Vec3<float>& wi, float& tmax, Vec3<float>& f,
                                           Vec3<float>& Le, float& wt, float& pl, bool& delta_out) {
  if (s.n_lights <= 0) return false;
  const float nl = float(s.n_lights);
  int li = int(ul * nl);
  li = li < s.n_lights - 1 ? li : s.n_lights - 1;
  Le = v3s(0.0f);
  pl = 0.0f;
  float dist = xinf<float>();
  bool delta = false;
  const int env_slot = s.has_sun ? 1 : 0;
  if (s.has_sun && li == 0) {
    wi = -v3f<float>(s.sun_dir);
    Le = v3f<float>(s.sun_E);
    pl = 1.0f;
    delta = true;
  } else if (s.env_type != 0 && li == env_slot) {
    env_sample<float>(s, ua, ub, wi, pl, Le);
  } else {
    const uint32_t code = s.emitters[li - env_slot - (s.env_type != 0 ? 1 : 0)];
    delta = sample_emitter<float>(s, code, x, ua, ub, wi, dist, pl, Le);
  }
  if (!(pl > 0.0f)) return false;
  float pf;
  if (MEDIUM) {
    const float p = hg_p(s.g, dot(d, wi));
    f = v3s(p);
    pf = p;
  } else {
    const Vec3<float> wol = to_lcs(wo, ns, ss, ts), wil = to_lcs(wi, ns, ss, ts);
    gm.eval_pdf(wol, wil, f, pf);
    f = f * fabsf(wil.z);
  }
  if (is_black(f)) return false;
  tmax = xisinf(dist) ? xinf<float>() : dist * float(1 - 1e-4);
  wt = delta ? 1.0f : power_heuristic(pl, pf);
  delta_out = delta;
  return true;
}

// Uniform in [0, 1) from (pixel, sample, vertex), independent of the path's
// own stream (the training reservoir draws no path random numbers, so a
// render samples the same paths with training on or off).
__device__ __forceinline__ float reservoir_u01(uint32_t pix, uint32_t smp, uint32_t depth) {
  uint32_t c[4] = {pix, smp, depth, 0x72657376u};
  philox10(c, 0x5bd1e995u, 0x1b873593u);
  return float(c[0] >> 8) * (1.0f / 16777216.0f);
}

// Guided mode: the learned proposal of every medium vertex that will be
// guided (depth < max_depth, below the bounce limit), 16 queries per MFMA tile,
// into gd.mix (40 floats per path slot).
__global__ __launch_bounds__(kChunk) void k_wf_guide(WfState w, WfArgs a, DevScene s, WfGuide gd) {
  __shared__ uint32_t wsum[kChunk / 64];
  __shared__ float gq[kChunk * 6];
  __shared__ uint32_t gslot[kChunk];
  __shared__ uint16_t gxa[4 * 16 * kLdA];
  __shared__ uint16_t gxb[4 * 16 * kLdA];
  __shared__ float gout[4 * 16 * kLdO];
  const WfPaths A = w.paths[a.cur];
  const uint32_t c = blockIdx.x, t = threadIdx.x, i = c * kChunk + t;
  bool q = false;
  if (t < chunk_paths(w, a, c) && A.kind[i] == WF_KIND_MEDIUM) {
    const int depth = int(A.depth[i] & 0xFFFFu);
    q = depth < gd.max_depth && depth < a.max_bounces;
  }
  uint32_t nq;
  const uint32_t qi = block_compact(q, wsum, nq);
  if (a.counters && t == 0 && nq) atomicAdd(&a.counters->guide_queries, (unsigned long long)nq);
  if (q) {
    const Vec3<float> o = v3(A.ox[i], A.oy[i], A.oz[i]), d = v3(A.dx[i], A.dy[i], A.dz[i]);
    const Vec3<float> p = xf_point<float>(s.inv, o + A.hit_t[i] * d);
    float* e = gq + 6 * qi;
    e[0] = p.x + 0.5f; e[1] = p.y + 0.5f; e[2] = p.z + 0.5f;
    e[3] = d.x; e[4] = d.y; e[5] = d.z;
    gslot[qi] = i;
  }
  __syncthreads();
  const int wv = int(t >> 6), lane = __lane_id();
  for (uint32_t tile = uint32_t(wv); tile * 16 < nq; tile += 4) {
    uint16_t* xa = gxa + wv * 16 * kLdA;
    uint16_t* xb = gxb + wv * 16 * kLdA;
    float* out = gout + wv * 16 * kLdO;
    wave_sync();
    guide_encode_tile(gd.net, xa, [&](int r, Vec3<float>& p, Vec3<float>& dd, bool& valid) {
      const uint32_t k = tile * 16 + uint32_t(r);
      valid = k < nq;
      const float* e = gq + 6 * (valid ? k : 0u);
      p = v3(e[0], e[1], e[2]);
      dd = v3(e[3], e[4], e[5]);
    });
    guide_mlp_tile(gd.net, xa, xb, out);
    if (lane < 16 && tile * 16 + uint32_t(lane) < nq)
      guide_head(out + lane * kLdO, gd.mix + 40 * size_t(gslot[tile * 16 + lane]));
  }
}

// Shading.  GUIDED: medium vertices at depth < max_depth sample a one-sample
// MIS mixture of HG and the learned vMF proposal (probabilities 1 - prob,
// prob) that k_wf_guide evaluated for them.  With
// training on, one guided vertex of each path (chosen uniformly by a
// reservoir) is recorded and becomes a training sample when the path ends:
// weight r = m Li * phase / pdf, Li the radiance gathered after the vertex
// divided by the throughput after it, m the path's number of guided vertices.
//
// SORT: before shading, the block orders its chunk's live paths by (event
// kind, material) with a bitonic sort of 256 keys in LDS, and thread t shades
// the path in sorted position t.  Waves then take one branch of the kind
// switch and one material's parameters; each path keeps its own RNG stream,
// so the image is unchanged.
//
// TRAIN (guided renders that train): the training record travels with the
// path; frozen guided renders carry none, and are held to the plain stage's
// 7 waves per SIMD (71 VGPRs, no spills).
// FUSE (volume-only scenes, plain renders): the stage tracks each vertex's
// shadow ray itself (ratio tracking, as k_wf_shadow does, 2: by 32-bit grid
// offsets, 1: 64-bit) and adds the light sample to the path's radiance or
// the film; no shadow queue, no shadow stage.
template <bool GUIDED, bool SORT, bool TRAIN, int FUSE = 0>
__global__ __launch_bounds__(kChunk, (GUIDED && !TRAIN) ? 7 : 1) void k_wf_shade(WfState w, WfArgs a, DevScene s, WfGuide gd) {
  __shared__ uint32_t wsum[kChunk / 64];
  const WfPaths A = w.paths[a.cur], B = w.paths[a.cur ^ 1];
  const WfShadow S = w.shadow;
  const uint32_t c = blockIdx.x, t = threadIdx.x;
  const uint32_t nlive = chunk_paths(w, a, c);
  if (nlive == 0) {  // an empty chunk (most of them once the pool drains): nothing continues
    if (t == 0) {
      w.cnt_paths[a.cur ^ 1][c] = 0u;
      w.cnt_shadow[c] = 0u;
      if (c == 0) w.status[1 + a.cur] = 0u;
    }
    return;
  }
  uint32_t slot = t;
  if constexpr (SORT) {
    __shared__ uint32_t skey[kChunk];
    uint32_t key = 0xFFFFFFu;  // free slots sort last
    if (t < nlive) {
      const uint32_t kd = A.kind[c * kChunk + t];
      uint32_t mat = 0;
      if (kd == WF_KIND_SURFACE) mat = uint32_t(hit_material(s, A.hit_ref[c * kChunk + t]));
      key = (kd << 21) | (mat < 0x1FFFFFu ? mat : 0x1FFFFFu);
    }
    skey[t] = (key << 8) | t;
    __syncthreads();
    for (uint32_t k = 2; k <= kChunk; k <<= 1)
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        const uint32_t p = t ^ j;
        if (p > t) {
          const uint32_t x0 = skey[t], x1 = skey[p];
          if ((x0 > x1) == ((t & k) == 0)) { skey[t] = x1; skey[p] = x0; }
        }
        __syncthreads();
      }
    slot = skey[t] & 0xFFu;
  }
  const uint32_t i = c * kChunk + slot;
  const bool active = t < nlive;
  bool shadow = false, cont = false, ends = false;
  Vec3<float> x = v3s(0.0f), nd = v3s(0.0f), beta = v3s(0.0f), wi = v3s(0.0f), f = v3s(0.0f), Le = v3s(0.0f);
  float tmax = 0.0f, wt = 0.0f, pl = 0.0f, ppdf = 0.0f;
  uint32_t pix = 0, smp = 0, rngn = 0, depth = 0, rngn0 = 0;
  float lx = 0.0f, ly = 0.0f, lz = 0.0f;
  // phase 1: what the segment hit
  uint32_t kind = WF_KIND_MISS;
  Vec3<float> o = v3s(0.0f), d = v3s(0.0f);
  bool vertex = false, medium = false;
  Glossy<float> gm;
  Vec3<float> ns = v3s(0.0f), ss = v3s(0.0f), ts = v3s(0.0f);
  // guided training record carried by the path
  uint32_t rst = 0;
  float rec[kGuideRecF];
  if (active) {
    kind = A.kind[i];
    o = v3(A.ox[i], A.oy[i], A.oz[i]);
    d = v3(A.dx[i], A.dy[i], A.dz[i]);
    const uint32_t dep = A.depth[i];
    depth = dep & 0xFFFFu;
    const bool prev_delta = (dep & kPrevDelta) != 0;
    pix = A.pix[i];
    smp = A.smp[i];
    lx = A.Lx[i]; ly = A.Ly[i]; lz = A.Lz[i];
    beta = v3(A.bx[i], A.by[i], A.bz[i]);
    if constexpr (TRAIN) {
      const WfGuideRec& R = gd.rec[a.cur];
      rst = R.state[i];
      if (rst != 0) {
        float* const src[kGuideRecF] = {R.px, R.py, R.pz, R.dx, R.dy, R.dz, R.wx, R.wy,
                                        R.wz, R.ratio, R.bx, R.by, R.bz, R.lx, R.ly, R.lz};
        for (int k = 0; k < kGuideRecF; ++k) rec[k] = src[k][i];
        if ((rst & 3u) == 1u) {  // radiance baseline: everything gathered before this segment
          rec[13] = lx; rec[14] = ly; rec[15] = lz;
          rst = (rst & ~3u) | 2u;
        }
      }
    }
    const float th = A.hit_t[i];
    // the draw count goes out with the record (the empty asm, after the
    // record's last load, keeps it from being sunk into the vertex branch,
    // one round trip later)
    rngn0 = A.rngn[i];
    asm volatile("" ::"v"(rngn0));
    if (kind == WF_KIND_MISS) {
      if (s.env_type != 0) {
        const float wm = prev_delta ? 1.0f : power_heuristic(A.prev_pdf[i], env_pdf<float>(s, d));
        const Vec3<float> E = env_le<float>(s, d);
        lx += beta.x * E.x * wm; ly += beta.y * E.y * wm; lz += beta.z * E.z * wm;
      }
      ends = true;
    } else if (kind == WF_KIND_SURFACE) {
      const uint32_t ref = A.hit_ref[i];
      SurfHit<float> hit;
      hit.t = th; hit.type = int(ref >> 30); hit.slot = ref & 0x3FFFFFFFu; hit.key = 0;
      Vec3<float> nrm;
      int mat = 0;
      hit_normal<float>(s, hit, o, d, nrm, mat);
      const float* m = s.mats + 12 * size_t(mat);
      if (m[0] == 1.0f) {  // emitter: MIS-weighted emission, the path ends
        const Vec3<float> E = v3f<float>(m + 6);
        float wm = 1.0f;
        if (!prev_delta) wm = power_heuristic(A.prev_pdf[i], emitter_hit_pdf<float>(s, hit.type, hit.slot, o, d, th, nrm));
        lx += beta.x * E.x * wm; ly += beta.y * E.y * wm; lz += beta.z * E.z * wm;
        ends = true;
      } else if (int(depth) >= a.max_bounces) {
        ends = true;
      } else {
        vertex = true;
        const Vec3<float> wo = -d;
        gm.albedo = v3f<float>(m + 1);
        const float rough = m[4];
        gm.alpha = rough * rough;
        gm.metallic = m[5];
        ns = dot(nrm, wo) < 0.0f ? -nrm : nrm;
        onb(ns, ss, ts);
      }
    } else {  // medium collision
      if (int(depth) >= a.max_bounces) {
        ends = true;
      } else {
        vertex = true;
        medium = true;
        beta = beta * v3(s.albedo[0], s.albedo[1], s.albedo[2]);
      }
    }
    if (vertex) x = o + th * d;
  }
  // phase 2: next-event estimation and the continuation
  if (active && vertex) {
    const Vec3<float> wo = -d;
    PathRng rng;
    rng.init(a.seed, pix, smp, a.job);
    rng.seek(rngn0);
    const float ul = rng.next(), ua = rng.next(), ub = rng.next();
    bool delta = false;
    shadow = medium ? nee_sample<true>(s, x, d, gm, wo, ns, ss, ts, ul, ua, ub, wi, tmax, f, Le, wt, pl, delta)
                    : nee_sample<false>(s, x, d, gm, wo, ns, ss, ts, ul, ua, ub, wi, tmax, f, Le, wt, pl, delta);
    const bool guided = GUIDED && medium && int(depth) < gd.max_depth;
    // the proposal: the tabulated entry of the vertex's cell and direction, or
    // the network's output for this path that k_wf_guide wrote
    const float* mix = nullptr;
    if (guided)
      mix = gd.cache.mix ? guide_cache_at(gd.cache, xf_point<float>(s.inv, x) + v3s(0.5f), d) : gd.mix + 40 * size_t(i);
    if (guided && shadow && !delta)  // the phase-sampling strategy is the mixture
      wt = power_heuristic(pl, (1.0f - gd.prob) * hg_p(s.g, dot(d, wi)) + gd.prob * vmf_mixture_pdf(mix, wi));
    if (shadow) f = beta * f * Le;  // NEE weight with the throughput before this vertex's BSDF weight
    if (medium && !guided) {
      const float u0 = rng.next(), u1 = rng.next();
      nd = hg_sample(s.g, d, u0, u1);
      ppdf = hg_p(s.g, dot(d, nd));
      cont = true;
    } else if (medium) {
      const float us = rng.next(), u0 = rng.next(), u1 = rng.next(), u2 = rng.next();
      nd = us < gd.prob ? vmf_mixture_sample(mix, u2, u0, u1) : hg_sample(s.g, d, u0, u1);
      const float ph = hg_p(s.g, dot(d, nd));
      const float pm = (1.0f - gd.prob) * ph + gd.prob * vmf_mixture_pdf(mix, nd);
      if (pm > 0.0f) {
        ppdf = pm;
        beta = beta * (ph / pm);
        cont = true;
        if constexpr (TRAIN) {
          // One record per path, a reservoir over its guided vertices: the
          // m-th guided vertex replaces the record with probability 1/m, so
          // every guided vertex is the recorded one with probability 1/m_end
          // (the sample's weight is scaled by m_end when the path ends).
          const uint32_t m = (rst >> 2) + 1u;
          const bool take = m == 1u || reservoir_u01(pix, smp, depth) * float(m) < 1.0f;
          if (take) {
            const Vec3<float> p = xf_point<float>(s.inv, x);
            rec[0] = p.x + 0.5f; rec[1] = p.y + 0.5f; rec[2] = p.z + 0.5f;
            rec[3] = d.x; rec[4] = d.y; rec[5] = d.z;
            rec[6] = nd.x; rec[7] = nd.y; rec[8] = nd.z;
            rec[9] = ph / pm;
            rec[10] = beta.x; rec[11] = beta.y; rec[12] = beta.z;
          }
          rst = (m << 2) | (take ? 1u : (rst & 3u));
        }
      }
    } else {
      const float u0 = rng.next(), u1 = rng.next(), u2 = rng.next();
      const Vec3<float> wol = to_lcs(wo, ns, ss, ts);
      const Vec3<float> wil = gm.sample(wol, u0, u1, u2);
      Vec3<float> fe;
      float pf;
      gm.eval_pdf(wol, wil, fe, pf);
      if (pf > 0.0f) {
        const Vec3<float> fs = fe * fabsf(wil.z);
        if (!is_black(fs)) {
          ppdf = pf;
          nd = to_world(wil, ns, ss, ts);
          cont = true;
          beta = beta * fs / pf;
        }
      }
    }
    ends = !cont;
    rngn = rng.consumed();
  }
  uint32_t ncont, nshadow = 0, k = 0;
  const uint32_t j = block_compact(cont, wsum, ncont);
  if constexpr (FUSE > 0) {
    // no queue
  } else if (a.shadow_full) {
    k = block_compact(shadow, wsum, nshadow);
  } else {
    // a continuing path's shadow ray sits at the path's own index j (a hole,
    // light pdf 0, where it has none); the rays of paths that end follow
    uint32_t nend;
    const uint32_t e = block_compact(shadow && !cont, wsum, nend);
    k = cont ? j : ncont + e;
    nshadow = ncont + nend;
  }
  if (cont) {
    const uint32_t b = c * kChunk + j;
    B.ox[b] = x.x; B.oy[b] = x.y; B.oz[b] = x.z;
    B.dx[b] = nd.x; B.dy[b] = nd.y; B.dz[b] = nd.z;
    B.bx[b] = beta.x; B.by[b] = beta.y; B.bz[b] = beta.z;
    if constexpr (FUSE == 0) { B.Lx[b] = lx; B.Ly[b] = ly; B.Lz[b] = lz; }  // else after the light sample
    B.pix[b] = pix; B.smp[b] = smp; B.rngn[b] = rngn;
    B.depth[b] = depth + 1u;  // prev_delta cleared
    B.prev_pdf[b] = ppdf;
    if constexpr (TRAIN) {
      const WfGuideRec& R = gd.rec[a.cur ^ 1];
      R.state[b] = rst;
      if (rst != 0) {
        float* const dst[kGuideRecF] = {R.px, R.py, R.pz, R.dx, R.dy, R.dz, R.wx, R.wy,
                                        R.wz, R.ratio, R.bx, R.by, R.bz, R.lx, R.ly, R.lz};
        for (int q = 0; q < kGuideRecF; ++q) dst[q][b] = rec[q];
      }
    }
  } else if (ends) {
    if (a.sbuf) sample_store(a, pix, smp, lx, ly, lz);
    else film_add(a.film_fixed, pix, lx, ly, lz);
  }
  if constexpr (TRAIN) {
    // a recording path that ends now becomes a training sample
    float r = 0.0f;
    if (ends && !cont && (rst & 3u) == 2u) {
      const float dl[3] = {lx - rec[13], ly - rec[14], lz - rec[15]};
      float li = 0.0f;
      int nc = 0;
      for (int q = 0; q < 3; ++q)
        if (rec[10 + q] > 0.0f) {
          li += dl[q] / rec[10 + q];
          ++nc;
        }
      r = nc ? li / float(nc) * rec[9] * float(rst >> 2) : 0.0f;
      if (!(r > 0.0f) || xisinf(r)) r = 0.0f;
    }
    uint32_t nsmp;
    const uint32_t si = block_compact(r > 0.0f, wsum, nsmp);
    __shared__ uint32_t sbase;
    if (t == 0) sbase = nsmp ? atomicAdd(gd.batch.count, nsmp) : 0u;
    __syncthreads();
    const uint32_t q = sbase + si;
    if (r > 0.0f && q < gd.batch.cap) {
      const uint32_t cap = gd.batch.cap;
      for (int e = 0; e < 9; ++e) gd.batch.s[size_t(e) * cap + q] = rec[e];
      gd.batch.s[size_t(9) * cap + q] = r;
    }
  }
  if constexpr (FUSE > 0) {
    // the shadow ray, tracked here: the draws and arithmetic of k_wf_shadow
    // and shadow_deliver, so a continuing path's radiance is the same one
    // addition (lx + c) and a final vertex's sample the same film atomic
    bool lit = false;
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    if (shadow) {
      PathRng prng;
      prng.init(a.seed, pix, smp, a.job);
      PathRng srng = prng.shadow_stream(int(depth));
      Count<false> nc;
      float tr = 1.0f;
      const Vec3<float> oo = xf_point<float>(s.inv, x), dd = xf_vec<float>(s.inv, wi);
      float ta, tb;
      if (box_interval<float>(s, oo, dd, ta, tb)) {
        ta = xmax(ta, 0.0f);
        tb = xmin(tb, tmax);
        if (ta < tb) tr = ratio_track<float, false, (FUSE == 2)>(s, x, wi, ta, tb, srng, nc);
      }
      if (tr > 0.0f) {
        const float q = tr * wt * float(s.n_lights) / pl;
        cx = f.x * q; cy = f.y * q; cz = f.z * q;
        lit = true;
        if (!cont) film_add(a.film_fixed, pix, cx, cy, cz);
      }
    }
    if (cont) {
      const uint32_t b = c * kChunk + j;
      B.Lx[b] = lit ? lx + cx : lx;
      B.Ly[b] = lit ? ly + cy : ly;
      B.Lz[b] = lit ? lz + cz : lz;
    }
  } else if (shadow) {
    const uint32_t q = c * kChunk + k;
    S.dx[q] = wi.x; S.dy[q] = wi.y; S.dz[q] = wi.z;
    S.tmax[q] = tmax;
    S.cx[q] = f.x; S.cy[q] = f.y; S.cz[q] = f.z;
    S.wt[q] = wt; S.pl[q] = pl;  // pl > 0: nee_sample has no ray otherwise
    if (!cont || a.shadow_full) {  // else the ray's index and the path's next record say these
      S.target[q] = cont ? j : (kToFilm | pix);
      S.ox[q] = x.x; S.oy[q] = x.y; S.oz[q] = x.z;
      S.pix[q] = pix; S.smp[q] = smp; S.vert[q] = depth;
    }
  } else if (cont && !a.shadow_full) {
    S.pl[c * kChunk + j] = 0.0f;  // a hole
  }
  if (t == 0) {
    w.cnt_paths[a.cur ^ 1][c] = ncont;
    w.cnt_shadow[c] = nshadow;
    // live flags: status[1 + b] != 0 iff buffer b holds a path; this launch
    // sets the next buffer's and clears its own (read, if at all, before it)
    if (ncont) w.status[2 - a.cur] = 1u;
    if (c == 0) w.status[1 + a.cur] = 0u;
  }
  if (a.counters) {
    uint32_t nend;
    (void)block_compact(ends && !cont, wsum, nend);
    if (t == 0 && nend) atomicAdd(&a.counters->paths, (unsigned long long)nend);
  }
}

// Shadow-queue layouts.  Full (WfArgs::shadow_full): the chunk's rays packed,
// each record whole.  Otherwise (scenes with a volume): slot k < ncont of the
// chunk is the ray of the continuing path at the same index of the next
// buffer (a hole, light pdf 0, if that path has none), whose record there
// holds the ray's origin (the vertex), pixel, sample and vertex (its depth
// minus one); the rays of paths that end follow, each record whole.
// own: slot k is a continuing path's.
// (Branch-free: each array chosen per lane, so the six loads go out together.)
__device__ __forceinline__ void shadow_source(const WfShadow& S, const WfPaths& B, uint32_t k, bool own,
                                              Vec3<float>& x, uint32_t& pix, uint32_t& smp, uint32_t& vert) {
  const SlotF ax = own ? B.ox : S.ox, ay = own ? B.oy : S.oy, az = own ? B.oz : S.oz;
  const SlotU ap = own ? B.pix : S.pix, as = own ? B.smp : S.smp, av = own ? B.depth : S.vert;
  x = v3(ax[k], ay[k], az[k]);
  pix = ap[k];
  smp = as[k];
  vert = av[k] - (own ? 1u : 0u);
}

// NEE contribution of an unoccluded shadow ray k (transmittance tr): to the
// continuing path in the next buffer (same chunk) or straight to the film.
__device__ __forceinline__ void shadow_deliver(const WfShadow& S, const WfPaths& B, const WfArgs& a,
                                               const DevScene& s, uint32_t k, float tr, bool own = false) {
  const float q = tr * S.wt[k] * float(s.n_lights) / S.pl[k];
  const float cx = S.cx[k] * q, cy = S.cy[k] * q, cz = S.cz[k] * q;
  if (own) {
    B.Lx[k] += cx; B.Ly[k] += cy; B.Lz[k] += cz;
    return;
  }
  const uint32_t tg = S.target[k];
  if (tg & kToFilm) {
    // the final vertex's light sample: a fire-and-forget atomic on the
    // fixed-point film (a read-modify-write of the sample slot would wait)
    film_add(a.film_fixed, tg & ~kToFilm, cx, cy, cz);
  } else {
    const uint32_t b = (k / kChunk) * kChunk + tg;
    B.Lx[b] += cx; B.Ly[b] += cy; B.Lz[b] += cz;
  }
}

// Shadow rays of scenes without a volume: any-hit BVH occlusion with lane
// refill over the block's pool of shadow-queue chunks, then delivery of the
// unoccluded rays' contributions.
template <bool C>
__global__ __launch_bounds__(kChunk) void k_wf_shadow_surf(WfState w, WfArgs a, DevScene s) {
  __shared__ uint32_t lds_stack[kShortStack * kChunk];
  __shared__ uint32_t red[kChunk / 64];
  __shared__ RefillPool pool;
  const WfShadow S = w.shadow;
  const WfPaths B = w.paths[a.cur ^ 1];
  pool_init(pool, w.cnt_shadow, a.nchunks);
  uint32_t* stk = lds_stack + threadIdx.x;
  uint32_t spill[kSpillStack];
  Count<C> cnt;
  TravState<float> st;
  uint32_t k = kNoHit;
  bool have = false;
  while (true) {
    if (refill_due(have, a.refill_min)) {
      const uint32_t got = pool_fetch(pool, !have);  // every lane still looping calls it
      if (!have) {
        if (got == kNoHit) break;
        k = got;
        have = true;
        cnt.shadow();
        trav_init(st, v3(S.ox[k], S.oy[k], S.oz[k]), v3(S.dx[k], S.dy[k], S.dz[k]), 1e-4f, S.tmax[k]);
      }
    }
    if (have && trav_step_wave<float, true, C>(s, st, stk, kChunk, spill, cnt)) {
      if (st.found) S.tmax[k] = -1.0f;  // occluded: marked for the delivery pass
      have = false;
    }
  }
  // Delivery pass over the block's chunks, one queue entry per thread in
  // order: coalesced, and out of the refill loop, where a finishing lane
  // would hold its whole wave through the delivery.
  __syncthreads();
  for (uint32_t j = 0; j < pool.n; ++j) {
    const uint32_t ch = pool_chunk(j), n = pool.pre[j + 1] - pool.pre[j];
    for (uint32_t t = threadIdx.x; t < n; t += kChunk) {
      const uint32_t q = ch * kChunk + t;
      if (S.tmax[q] >= 0.0f) shadow_deliver(S, B, a, s, q, 1.0f);
    }
  }
  if constexpr (C) {
    block_count<C>(&a.counters->shadow_rays, cnt.n_shadow, red);
    block_count<C>(&a.counters->sh_nodes, cnt.n_node, red);
    block_count<C>(&a.counters->sh_prims, cnt.n_prim, red);
  }
}

// Shadow rays: any-hit occlusion (HAS_SURF) then ratio tracking (HAS_VOL),
// the megakernel's transmittance() split by scene content.
// SPATIAL: the chunk's shadow rays are taken in Morton order of their origins
// (each keeps its own stream: same transmittances).
// SMALL: the scene's grid is under 4 GiB (density lookups by 32-bit offsets).
template <bool C, bool HAS_SURF, bool HAS_VOL, bool SPATIAL, bool SMALL>
__global__ __launch_bounds__(kChunk) void k_wf_shadow(WfState w, WfArgs a, DevScene s) {
  __shared__ uint32_t lds_stack[HAS_SURF ? kShortStack * kChunk : 1];
  __shared__ uint32_t red[kChunk / 64];
  __shared__ uint32_t skey[SPATIAL ? kChunk : 1];
  const WfShadow S = w.shadow;
  const WfPaths B = w.paths[a.cur ^ 1];
  const uint32_t c = blockIdx.x, t = threadIdx.x;
  const uint32_t n = w.cnt_shadow[c];
  const uint32_t nc = a.shadow_full ? 0u : w.cnt_paths[a.cur ^ 1][c];  // slots of continuing paths' rays
  uint32_t slot = t;
  if constexpr (SPATIAL) {
    const uint32_t q = c * kChunk + t;
    slot = chunk_spatial_slot(s, t < n,
                              t < n ? (t < nc ? v3(B.ox[q], B.oy[q], B.oz[q]) : v3(S.ox[q], S.oy[q], S.oz[q]))
                                    : v3s(0.0f),
                              skey);
  }
  const uint32_t k = c * kChunk + slot;
  const bool own = slot < nc;
  Count<C> cnt;
  Vec3<float> x = v3s(0.0f), wi = v3s(0.0f);
  uint32_t spix = 0, ssmp = 0, svert = 0;
  float tmax = 0.0f, pl = 0.0f;
  // every load of the ray goes out at once, before the hole test (a hole's
  // record and path record exist, so the loads stay in bounds); the empty
  // asm takes the values here, so the loads are not sunk past the test into
  // one round trip each
  if (t < n) {
    shadow_source(S, B, k, own, x, spix, ssmp, svert);
    wi = v3(S.dx[k], S.dy[k], S.dz[k]);
    tmax = S.tmax[k];
    pl = S.pl[k];
    asm volatile("" ::"v"(x.x), "v"(x.y), "v"(x.z), "v"(spix), "v"(ssmp), "v"(svert), "v"(wi.x), "v"(wi.y),
                 "v"(wi.z), "v"(tmax), "v"(pl));
  }
  if (t < n && (!own || pl > 0.0f)) {  // holes skipped
    PathRng rng;
    rng.init(a.seed, spix, ssmp, a.job);
    PathRng srng = rng.shadow_stream(int(svert));
    cnt.shadow();
    float tr = 1.0f;
    if (HAS_SURF) {
      SurfHit<float> h;
      if (traverse<float, true, C>(s, x, wi, 1e-4f, tmax, lds_stack + t, kChunk, h, cnt)) tr = 0.0f;
    }
    if (HAS_VOL && tr > 0.0f) {
      const Vec3<float> oo = xf_point<float>(s.inv, x), dd = xf_vec<float>(s.inv, wi);
      float ta, tb;
      if (box_interval<float>(s, oo, dd, ta, tb)) {
        ta = xmax(ta, 0.0f);
        tb = xmin(tb, tmax);
        if (ta < tb) tr = ratio_track<float, C, SMALL>(s, x, wi, ta, tb, srng, cnt);
      }
    }
    if (tr > 0.0f) shadow_deliver(S, B, a, s, k, tr, own);
  }
  if constexpr (C) {
    block_count<C>(&a.counters->shadow_rays, cnt.n_shadow, red);
    block_count<C>(&a.counters->sh_lookups, cnt.n_lookup, red);
    block_count<C>(&a.counters->sh_cells, cnt.n_cell, red);
    block_count<C>(&a.counters->sh_nodes, cnt.n_node, red);
    block_count<C>(&a.counters->sh_prims, cnt.n_prim, red);
  }
}

// film += fixed / 2^30 and clear the fixed-point buffer
// film pixel += sum of its samples (FP64, in sample order) + its fixed-point
// sum, which is cleared
__global__ void k_wf_flush_samples(const float* sbuf, long long* fixed, double* film, int W, int x0, int y0, int rw,
                                   int rh, uint32_t spp) {
  const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x, npix = size_t(rw) * size_t(rh);
  if (p >= npix) return;
  double r = 0.0, g = 0.0, b = 0.0;
  for (uint32_t s = 0; s < spp; ++s) {
    const float* q = sbuf + 3 * (size_t(s) * npix + p);
    r += double(q[0]);
    g += double(q[1]);
    b += double(q[2]);
  }
  const size_t pix = (size_t(y0) + p / size_t(rw)) * size_t(W) + size_t(x0) + p % size_t(rw);
  long long* fx = fixed + 3 * pix;
  const double inv = 1.0 / 1073741824.0;
  film[3 * pix] += r + double(fx[0]) * inv;
  film[3 * pix + 1] += g + double(fx[1]) * inv;
  film[3 * pix + 2] += b + double(fx[2]) * inv;
  fx[0] = 0;
  fx[1] = 0;
  fx[2] = 0;
}

// Batch films keep 2^-30 fixed-point sums: each pixel's samples are summed
// as above and the sum, rounded once, joins the pixel's fixed-point sum.
__global__ void k_wf_flush_samples_fixed(const float* sbuf, long long* fixed, int W, int x0, int y0, int rw, int rh,
                                         uint32_t spp) {
  const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x, npix = size_t(rw) * size_t(rh);
  if (p >= npix) return;
  double r = 0.0, g = 0.0, b = 0.0;
  for (uint32_t s = 0; s < spp; ++s) {
    const float* q = sbuf + 3 * (size_t(s) * npix + p);
    r += double(q[0]);
    g += double(q[1]);
    b += double(q[2]);
  }
  const size_t pix = (size_t(y0) + p / size_t(rw)) * size_t(W) + size_t(x0) + p % size_t(rw);
  long long* fx = fixed + 3 * pix;
  fx[0] += __double2ll_rn(r * 1073741824.0);
  fx[1] += __double2ll_rn(g * 1073741824.0);
  fx[2] += __double2ll_rn(b * 1073741824.0);
}

__global__ void k_wf_flush_film(long long* fixed, double* film, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long v = fixed[i];
  if (v != 0) {
    film[i] += double(v) * (1.0 / 1073741824.0);
    fixed[i] = 0;
  }
}

// ---- pool layout ------------------------------------------------------------
// A and B path buffers (15 float + 6 uint32 arrays each), the shadow queue
// (12 float + 4 uint32 arrays), per-chunk counters (A, B, shadow, regen count
// and 64-bit regen base) and the status / work-counter words, each array
// 256-byte aligned, in one allocation.
static constexpr int kPathF = 15, kPathU = 6, kShadowF = 12, kShadowU = 4;
static inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

size_t wf_bytes(uint32_t P) {
  const size_t NC = P / kChunk, NB = (NC + kPlanThreads - 1) / kPlanThreads;
  return size_t(2 * (kPathF + kPathU) + kShadowF + kShadowU) * align256(size_t(P) * 4) + 4 * align256(NC * 4) +
         align256(NC * 8) + 2 * 256 + align256(NB * 8);
}

void wf_carve(void* base, uint32_t P, WfState& w) {
  char* p = static_cast<char*>(base);
  const size_t stride = align256(size_t(P) * 4);
  const size_t NC = P / kChunk;
  for (int b = 0; b < 2; ++b) {
    WfPaths& q = w.paths[b];
    SlotF* fl[kPathF] = {&q.ox, &q.oy, &q.oz, &q.dx, &q.dy, &q.dz, &q.bx, &q.by,
                         &q.bz, &q.Lx, &q.Ly, &q.Lz, &q.prev_pdf, &q.hit_t, &q.pad};
    for (int k = 0; k < kPathF; ++k, p += stride) fl[k]->p = reinterpret_cast<float*>(p);
    SlotU* us[kPathU] = {&q.pix, &q.smp, &q.rngn, &q.depth, &q.hit_ref, &q.kind};
    for (int k = 0; k < kPathU; ++k, p += stride) us[k]->p = reinterpret_cast<uint32_t*>(p);
  }
  WfShadow& S = w.shadow;
  SlotF* sf[kShadowF] = {&S.ox, &S.oy, &S.oz, &S.dx, &S.dy, &S.dz, &S.tmax, &S.cx, &S.cy, &S.cz, &S.wt, &S.pl};
  for (int k = 0; k < kShadowF; ++k, p += stride) sf[k]->p = reinterpret_cast<float*>(p);
  SlotU* su[kShadowU] = {&S.target, &S.pix, &S.smp, &S.vert};
  for (int k = 0; k < kShadowU; ++k, p += stride) su[k]->p = reinterpret_cast<uint32_t*>(p);
  uint32_t** cs[4] = {&w.cnt_paths[0], &w.cnt_paths[1], &w.cnt_shadow, &w.regen_m};
  for (int k = 0; k < 4; ++k, p += align256(NC * 4)) *cs[k] = reinterpret_cast<uint32_t*>(p);
  w.regen_base = reinterpret_cast<unsigned long long*>(p);
  p += align256(NC * 8);
  w.status = reinterpret_cast<uint32_t*>(p);
  p += 256;
  w.next_path = reinterpret_cast<unsigned long long*>(p);
  p += 256;
  w.plan_lb = reinterpret_cast<unsigned long long*>(p);
}

// guided-mode arrays: 2 buffers x (kGuideRecF floats + 1 state word) per slot,
// then 40 mixture floats per slot
size_t wf_guide_bytes(uint32_t P) {
  return size_t(2 * (kGuideRecF + kGuideRecU)) * align256(size_t(P) * 4) + align256(size_t(P) * 40 * 4);
}

void wf_guide_carve(void* base, uint32_t P, WfGuide& g) {
  char* p = static_cast<char*>(base);
  const size_t stride = align256(size_t(P) * 4);
  for (int b = 0; b < 2; ++b) {
    WfGuideRec& R = g.rec[b];
    float** fl[kGuideRecF] = {&R.px, &R.py, &R.pz, &R.dx, &R.dy, &R.dz, &R.wx, &R.wy,
                              &R.wz, &R.ratio, &R.bx, &R.by, &R.bz, &R.lx, &R.ly, &R.lz};
    for (int k = 0; k < kGuideRecF; ++k, p += stride) *fl[k] = reinterpret_cast<float*>(p);
    R.state = reinterpret_cast<uint32_t*>(p);
    p += stride;
  }
  g.mix = reinterpret_cast<float*>(p);
}

__global__ void k_guide_reset_count(uint32_t* count) { *count = 0u; }

hipError_t wf_refill_blocks_per_cu(int* blocks) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, reinterpret_cast<const void*>(&k_wf_trace<false>),
                                                      int(kChunk), 0);
}

hipError_t launch_wavefront(const DevScene& s, const DevCamera& cam, WfState& w, WfArgs a, bool count,
                            hipStream_t st, uint32_t* h_status, StageTimer* tm, WfGuideHook* gh) {
  hipError_t e = hipSuccess;
#define NE_MARK(stage)                                          \
  do {                                                          \
    if (tm && (e = tm->mark(stage)) != hipSuccess) return e;    \
  } while (0)
  const uint32_t NC = a.nchunks;
  NE_MARK(kStageOther);
  hipLaunchKernelGGL(k_wf_init, dim3((NC + 255) / 256), dim3(256), 0, st, w, NC);
  const dim3 grid = dim3(NC);
  const dim3 pgrid = dim3((NC + a.refill - 1) / a.refill);  // lane-refill kernels
  const dim3 blk = dim3(unsigned(kChunk));
  if (!count) a.counters = nullptr;
  const bool hs = s.n_nodes > 0, hv = s.has_volume != 0;
  const bool sort = (a.sort & 1) != 0;
  const bool spatial = (a.sort & 2) != 0;  // tracking stages in Morton order of the ray origins
  const bool small = s.grid_small != 0 && !a.wide_offsets;  // density lookups by 32-bit offsets
  // the volume shadow stage reads a continuing path's origin and RNG key
  // from its next record; the surface-only shadow stage reads whole records
  a.shadow_full = hv ? 0 : 1;
  // Drain: once every path of the render has been issued, the plan would
  // only count the paths in flight, so it is no longer launched (nor regen);
  // the loop ends on the live flag the shading stage keeps instead.
  bool drain = false;
  for (uint32_t it = 0;; ++it) {
    a.cur = int(it & 1u);
    const bool planned = !drain;  // this iteration's new paths are planned
    if (planned) {
      NE_MARK(kStagePlan);
      a.plan_epoch += 1u;
      hipLaunchKernelGGL(k_wf_plan, dim3((NC + kPlanThreads - 1) / kPlanThreads), dim3(unsigned(kPlanThreads)), 0,
                         st, w, a);
    }
    // every 8 iterations (every iteration while the guide trains): stop once
    // nothing is in flight (status written by plan, or the live flag);
    // guided training: a step on the queued samples when enough are in
    const bool training = gh && gh->g.train;
    if ((it & 7u) == 7u || (training && it > 0)) {
      if ((e = hipMemcpyAsync(h_status, w.status + (planned ? 0 : 1 + a.cur), sizeof(uint32_t),
                              hipMemcpyDeviceToHost, st)) != hipSuccess)
        return e;
      if (planned && (e = hipMemcpyAsync(h_status + 2, w.next_path, sizeof(unsigned long long),
                                         hipMemcpyDeviceToHost, st)) != hipSuccess)
        return e;
      if ((e = hipMemcpyAsync(h_status + 4, w.status + 3, sizeof(uint32_t), hipMemcpyDeviceToHost, st)) !=
          hipSuccess)
        return e;
      if (gh && gh->g.train &&
          (e = hipMemcpyAsync(h_status + 1, gh->g.batch.count, sizeof(uint32_t), hipMemcpyDeviceToHost, st)) !=
              hipSuccess)
        return e;
      if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
      if (h_status[4] != 0) return hipErrorLaunchTimeOut;  // a plan block's look-back gave up
      if (h_status[0] == 0) break;
      if (planned) {
        unsigned long long issued;
        std::memcpy(&issued, h_status + 2, sizeof(issued));
        drain = issued >= a.total_paths;  // from the next iteration on
      }
      if (gh && gh->g.train && h_status[1] >= gh->min_batch) {
        NE_MARK(kStageTrain);
        if ((e = launch_guide_grad(gh->g.net, gh->g.batch, gh->train, st)) != hipSuccess) return e;
        gh->steps += 1;
        gh->samples += std::min(h_status[1], gh->g.batch.cap);
        if ((e = launch_guide_adam(gh->g.net, gh->lr_table, gh->lr_mlp, gh->beta1, gh->beta2, gh->eps,
                                   int(gh->steps), st)) != hipSuccess)
          return e;
        if (gh->cache_mix && (e = launch_guide_bake(gh->g.net, gh->g.cache, gh->cache_mix, st)) != hipSuccess)
          return e;
        hipLaunchKernelGGL(k_guide_reset_count, dim3(1), dim3(1), 0, st, gh->g.batch.count);
      }
    }
    const bool fused = hv && !hs && !spatial && planned;  // regen inside the medium stage
    a.regen_fused = fused ? 1 : 0;
    if (!fused && planned) {
      NE_MARK(kStageRegen);
      hipLaunchKernelGGL(k_wf_regen, grid, blk, 0, st, w, a, cam);
    }
    if (hs) {
      NE_MARK(kStageTrace);
      if (count) hipLaunchKernelGGL(k_wf_trace<true>, pgrid, blk, 0, st, w, a, s);
      else hipLaunchKernelGGL(k_wf_trace<false>, pgrid, blk, 0, st, w, a, s);
    }
    if (hv) {
      NE_MARK(kStageMedium);
#define NE_MEDIUM(CNT, S_, SP, RG)                                                                       \
  do {                                                                                                   \
    if (small) hipLaunchKernelGGL((k_wf_medium<CNT, S_, SP, RG, true>), grid, blk, 0, st, w, a, s, cam); \
    else hipLaunchKernelGGL((k_wf_medium<CNT, S_, SP, RG, false>), grid, blk, 0, st, w, a, s, cam);     \
  } while (0)
      if (spatial) {
        if (count) { if (hs) NE_MEDIUM(true, true, true, false); else NE_MEDIUM(true, false, true, false); }
        else { if (hs) NE_MEDIUM(false, true, true, false); else NE_MEDIUM(false, false, true, false); }
      } else {
        if (count) {
          if (hs) NE_MEDIUM(true, true, false, false);
          else if (fused) NE_MEDIUM(true, false, false, true);
          else NE_MEDIUM(true, false, false, false);
        } else {
          if (hs) NE_MEDIUM(false, true, false, false);
          else if (fused) NE_MEDIUM(false, false, false, true);
          else NE_MEDIUM(false, false, false, false);
        }
      }
#undef NE_MEDIUM
    }
    if (!hs && !hv) {
      NE_MARK(kStageOther);
      hipLaunchKernelGGL(k_wf_all_miss, grid, blk, 0, st, w, a);
    }
    if (gh && !gh->g.cache.mix) {  // per-vertex network evaluation (no tabulated proposal)
      NE_MARK(kStageGuide);
      hipLaunchKernelGGL(k_wf_guide, grid, blk, 0, st, w, a, s, gh->g);
    }
    NE_MARK(kStageShade);
    // volume-only renders (plain, or with a frozen guide and a grid under
    // 4 GiB: the other guided variant would spill): shade tracks the shadow
    // rays itself
    const bool fuse = hv && !hs && !count && !spatial && !training && (!gh || small);
#define NE_SHADE_FUSED(G, SO, GD)                                                                           \
  do {                                                                                                      \
    if (small) hipLaunchKernelGGL((k_wf_shade<G, SO, false, 2>), grid, blk, 0, st, w, a, s, GD);           \
    else hipLaunchKernelGGL((k_wf_shade<G, SO, false, 1>), grid, blk, 0, st, w, a, s, GD);                 \
  } while (0)
    if (fuse && gh && sort) NE_SHADE_FUSED(true, true, gh->g);
    else if (fuse && gh) NE_SHADE_FUSED(true, false, gh->g);
    else if (fuse && sort) NE_SHADE_FUSED(false, true, WfGuide{});
    else if (fuse) NE_SHADE_FUSED(false, false, WfGuide{});
#undef NE_SHADE_FUSED
    else if (gh && training && sort) hipLaunchKernelGGL((k_wf_shade<true, true, true>), grid, blk, 0, st, w, a, s, gh->g);
    else if (gh && training) hipLaunchKernelGGL((k_wf_shade<true, false, true>), grid, blk, 0, st, w, a, s, gh->g);
    else if (gh && sort) hipLaunchKernelGGL((k_wf_shade<true, true, false>), grid, blk, 0, st, w, a, s, gh->g);
    else if (gh) hipLaunchKernelGGL((k_wf_shade<true, false, false>), grid, blk, 0, st, w, a, s, gh->g);
    else if (sort) hipLaunchKernelGGL((k_wf_shade<false, true, false>), grid, blk, 0, st, w, a, s, WfGuide{});
    else hipLaunchKernelGGL((k_wf_shade<false, false, false>), grid, blk, 0, st, w, a, s, WfGuide{});
    if (!fuse) NE_MARK(kStageShadow);
#define NE_SHADOW_SP(CNT, S_, V_, SM)                                                             \
  do {                                                                                            \
    if (spatial) hipLaunchKernelGGL((k_wf_shadow<CNT, S_, V_, true, SM>), grid, blk, 0, st, w, a, s); \
    else hipLaunchKernelGGL((k_wf_shadow<CNT, S_, V_, false, SM>), grid, blk, 0, st, w, a, s);      \
  } while (0)
#define NE_SHADOW(CNT, S_, V_)                       \
  do {                                               \
    if (V_ && small) NE_SHADOW_SP(CNT, S_, V_, true); \
    else NE_SHADOW_SP(CNT, S_, V_, false);           \
  } while (0)
    if (fuse) {
      // done in the shading stage
    } else if (hs && !hv) {
      if (count) hipLaunchKernelGGL(k_wf_shadow_surf<true>, pgrid, blk, 0, st, w, a, s);
      else hipLaunchKernelGGL(k_wf_shadow_surf<false>, pgrid, blk, 0, st, w, a, s);
    } else if (count) {
      if (hs) NE_SHADOW(true, true, true);
      else if (hv) NE_SHADOW(true, false, true); else NE_SHADOW(true, false, false);
    } else {
      if (hs) NE_SHADOW(false, true, true);
      else if (hv) NE_SHADOW(false, false, true); else NE_SHADOW(false, false, false);
    }
#undef NE_SHADOW
#undef NE_SHADOW_SP
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  const size_t nf = size_t(a.W) * size_t(a.H) * 3;
  NE_MARK(kStageOther);
  if (a.sbuf) {
    const int rw = a.x1 - a.x0, rh = a.y1 - a.y0;
    const size_t npix = size_t(rw) * size_t(rh);
    const uint32_t spp = uint32_t(a.total_paths / npix);
    if (a.film)
      hipLaunchKernelGGL(k_wf_flush_samples, dim3(uint32_t((npix + 255) / 256)), dim3(256), 0, st, a.sbuf,
                         a.film_fixed, a.film, a.W, a.x0, a.y0, rw, rh, spp);
    else  // the caller keeps the fixed-point sums (batch films)
      hipLaunchKernelGGL(k_wf_flush_samples_fixed, dim3(uint32_t((npix + 255) / 256)), dim3(256), 0, st, a.sbuf,
                         a.film_fixed, a.W, a.x0, a.y0, rw, rh, spp);
  } else if (a.film) {  // a.film == nullptr: the caller keeps the fixed-point sums (batch films)
    hipLaunchKernelGGL(k_wf_flush_film, dim3(uint32_t((nf + 255) / 256)), dim3(256), 0, st, a.film_fixed, a.film, nf);
  }
  NE_MARK(kStageEnd);
#undef NE_MARK
  return hipGetLastError();
}

}  // namespace ne
