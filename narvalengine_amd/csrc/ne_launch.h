// ne_launch.h -- host-callable launchers of the HIP kernels (C++ linkage,
// internal to the library; the public surface is include/ne_gpu.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "ne_scene.h"
#include "ne_vmf.h"

namespace ne {

// Per-launch timing: an event is recorded before every kernel launch and once
// at the end; the interval from one event to the next is charged to the stage
// of the launch between them.  Events come from a pool that grows on demand.
struct StageTimer {
  std::vector<hipEvent_t>* pool;
  std::vector<int8_t>* stages;
  hipStream_t st;
  hipError_t mark(int stage) {
    const size_t k = stages->size();
    if (k == pool->size()) {
      hipEvent_t e;
      const hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      pool->push_back(e);
    }
    stages->push_back(int8_t(stage));
    return hipEventRecord((*pool)[k], st);
  }
};

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
  GuideUse guide;   // guide.cache.mix != nullptr: guided medium vertices (FP32 only)
};

// Whether a scene needs the surface code of the path tracers: a BVH or any
// emitter (a volume lit only by the sun and the environment does not).
inline bool scene_has_surfaces(const DevScene& s) {
  const int emitters = s.n_lights - (s.has_sun ? 1 : 0) - (s.env_type != 0 ? 1 : 0);
  return s.n_nodes > 0 || emitters > 0;
}

hipError_t launch_render(const DevScene& s, const DevCamera& cam, const RenderArgs& a, int precision, bool count,
                         int block, int occupancy, hipStream_t st);

// ---- wavefront pipeline ---------------------------------------------------
// The path pool is split into chunks of kChunk slots; stage kernels run one
// block of kChunk threads per chunk and keep each chunk's live paths packed.
constexpr uint32_t kChunk = 256;
constexpr uint32_t kMaxRefill = 64;  // chunks per block in the lane-refill kernels
// Idle lanes a wave waits for before it refills.  Measured on MI355X with
// surface1m (ms per frame): 1: 444, 8: 427, 16: 421, 24: 416, 32: 417,
// 48: 459, 64: 771.
constexpr int kRefillMinIdle = 24;
// stage ids (the public NE_STAGE_* values)
enum { kStageMegakernel = 0, kStagePlan, kStageRegen, kStageTrace, kStageMedium, kStageShade, kStageShadow,
       kStageOther, kStageGuide, kStageTrain, kStageCount, kStageEnd = -1 };
enum { WF_KIND_MEDIUM = 0, WF_KIND_SURFACE = 1, WF_KIND_MISS = 2 };
constexpr uint32_t kNoHit = 0xFFFFFFFFu;
constexpr uint32_t kPrevDelta = 1u << 16;  // depth word: set while the last vertex was the camera

// A per-slot array of 4-byte elements.  Pools hold at most 2^28 slots, so
// element i lies 4 * i < 2^32 bytes from the base: indexing by that 32-bit
// byte offset lets a device access take the scalar-base form (a uniform
// base, one 32-bit lane offset shared by every array) instead of a 64-bit
// address per lane and array.
template <typename T>
struct SlotArray {
  static_assert(sizeof(T) == 4, "4-byte elements");
  T* p;
  __host__ __device__ T& operator[](uint32_t i) const {
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + size_t(i * 4u));
  }
};
using SlotF = SlotArray<float>;
using SlotU = SlotArray<uint32_t>;

// One path buffer, structure of arrays over P slots.
struct WfPaths {
  SlotF ox, oy, oz, dx, dy, dz;  // current ray
  SlotF bx, by, bz;              // path throughput
  SlotF Lx, Ly, Lz;              // radiance gathered so far
  SlotF prev_pdf;                // pdf of the sampled direction (emission MIS)
  SlotF hit_t;                   // surface hit / medium collision distance
  SlotF pad;
  SlotU pix, smp, rngn, depth;   // pixel, sample, draws consumed, depth | kPrevDelta
  SlotU hit_ref;                 // (prim type << 30) | storage slot
  SlotU kind;                    // WF_KIND_*
};
// Shadow-ray queue (per chunk, packed): origin, direction, extent, (beta f) Le,
// MIS weight, light pdf, target (slot within the chunk of the next buffer, or
// film pixel | 2^31), RNG key.  Surface-only scenes pack each chunk's rays
// with whole records (WfArgs::shadow_full).  Scenes with a volume put a
// continuing path's ray at the path's own index in the chunk (light pdf 0
// where it has none), with no target, origin or RNG key (the path's next
// record holds them), and the rays of paths that end after those, whole.
struct WfShadow {
  SlotF ox, oy, oz, dx, dy, dz, tmax, cx, cy, cz, wt, pl;
  SlotU target, pix, smp, vert;
};
struct WfState {
  WfPaths paths[2];
  WfShadow shadow;
  uint32_t* cnt_paths[2];          // live paths per chunk, per buffer
  uint32_t* cnt_shadow;            // shadow rays per chunk
  uint32_t* regen_m;               // new paths per chunk this iteration
  unsigned long long* regen_base;  // first new path id per chunk
  uint32_t* status;                // [0]: paths in flight this iteration; [1], [2]: live flags of the
                                   // two path buffers; [3]: nonzero if a plan look-back gave up
  unsigned long long* next_path;   // work counter
  unsigned long long* plan_lb;     // plan look-back words, one per plan block
};
size_t wf_bytes(uint32_t P);
// resident 256-thread blocks per CU of the lane-refill trace kernel
hipError_t wf_refill_blocks_per_cu(int* blocks);
void wf_carve(void* base, uint32_t P, WfState& w);

struct WfArgs {
  int32_t W, H, x0, y0, x1, y1;
  int32_t max_bounces;
  uint32_t sample_offset, job, nchunks;
  uint32_t refill;  // chunks per block of the lane-refill kernels, 1..kMaxRefill
  uint64_t seed;
  unsigned long long total_paths;
  int32_t cur;
  DevCounters* counters;
  long long* film_fixed;
  float* sbuf;  // per-sample radiance (3 floats per path of the render), or null: fixed-point film atomics
  double* film;
  uint32_t* gstate[2];  // guided mode: training-record state per path buffer (else null)
  int32_t sort;         // bit 0: shade sorts each chunk by (event kind, material);
                        // bit 1: the tracking stages take each chunk's rays in Morton order
  int32_t refill_min;   // refill kernels: a wave refills once this many of its lanes are idle
  uint32_t plan_epoch;  // tags this plan launch's look-back words (1, 2, ... within a render)
  int32_t shadow_full;  // every shadow record carries its origin and RNG key (surface-only scenes)
  int32_t wide_offsets;  // density lookups by 64-bit offsets whatever the grid's size
  uint32_t plan_stall;   // test hook: the plan's look-back waits for an epoch that never comes
  int32_t regen_fused;  // this iteration's medium stage makes the new paths and leaves the live
                        // counts as they were: later stages add regen_m
};

struct WfGuideHook;  // guided mode: proposal, training records, online training
// Renders a.total_paths paths through a pool of a.nchunks * kChunk slots;
// returns after the film flush is enqueued.  a.film == nullptr: the sums stay
// in a.film_fixed (2^-30 fixed point; per-sample slots, if any, are summed into
// it).  h_status: pinned host words.
hipError_t launch_wavefront(const DevScene& s, const DevCamera& cam, WfState& w, WfArgs a, bool count,
                            hipStream_t st, uint32_t* h_status, StageTimer* tm, WfGuideHook* gh);
// ---- persistent batched renderer -------------------------------------------
constexpr int kMaxBatchJobs = 1024;
struct BatchJob {
  DevScene scene;
  DevCamera cam;
  int32_t W, H, spp, max_bounces;
  uint64_t seed;
  uint32_t job;
  uint32_t tiles_x, tiles_y, chunks;  // work items of the job: tiles_x * tiles_y * chunks
  unsigned long long first_item;      // global index of its first item
  long long* film;                    // W * H * 3 fixed-point sums
  GuideUse guide;                     // guide.cache.mix != nullptr: guided medium vertices
};
struct BatchArgs {
  const BatchJob* jobs;
  int32_t njobs;
  uint32_t tile_w, tile_h, spi;       // item shape: tile and samples per item
  unsigned long long total_items;
  unsigned long long* next_item;      // queue counter
  unsigned long long* item_time;      // per-item completion timestamp (wall clock ticks)
  unsigned long long* start_time;
  float* job_ms;                      // per-job completion time
  long long* films;                   // all films, contiguous
};
// surfaces = false: every job's scene is a volume lit only by the sun and the
// environment (no BVH, no emitters); the kernel then carries no surface code.
// guided: some job samples a guide's tabulated proposal.
hipError_t batch_blocks_per_cu(bool surfaces, bool guided, int* blocks);
// init (clear films, queue, timestamps) + persistent kernel + per-job times
hipError_t launch_batch(const BatchArgs& b, bool surfaces, bool guided, size_t film_words, uint32_t grid_blocks,
                        double ticks_per_ms, hipStream_t st);
hipError_t launch_fixed_to_double(const long long* in, double* out, size_t n, hipStream_t st);

// ---- learned directional guide ------------------------------------------------
// Hash-grid encoding (kGuideLevels x 2 features) + MLP 35 -> 64 x 4 (ReLU) -> 40
// (8 vMF lobes x {logit, mean xyz, log kappa}), widths padded to multiples of
// 16 for the 16x16x16 BF16 MFMA tiles: layer l maps kGuideIn[l] -> kGuideOut[l].
constexpr int kGuideLevels = 16, kGuideFeat = 2, kGuideLobes = kVmfLobes, kGuideLayers = 5;
constexpr int kGuideEnc = kGuideLevels * kGuideFeat;  // 32
constexpr int kGuideInputs = kGuideEnc + 3;           // 35: encoding + direction
constexpr int kGuideOutputs = kGuideLobes * 5;        // 40
constexpr int kGuideWidth = 64;                       // activation row width in memory
__host__ __device__ constexpr int guide_in(int l) { return l == 0 ? 48 : 64; }
__host__ __device__ constexpr int guide_out(int l) { return l == kGuideLayers - 1 ? 48 : 64; }
__host__ __device__ constexpr int guide_w_off(int l) {
  return l == 0 ? 0 : guide_w_off(l - 1) + guide_in(l - 1) * guide_out(l - 1);
}
constexpr int kGuideWeights = guide_w_off(kGuideLayers);                                  // 18432
__host__ __device__ constexpr int guide_b_off(int l) {
  return kGuideWeights + (l == 0 ? 0 : guide_b_off(l - 1) - kGuideWeights + guide_out(l - 1));
}
constexpr int kGuideParams = guide_b_off(kGuideLayers);                                   // + 304 biases

struct GuideNet {
  float* table;        // kGuideLevels * T * 2 hash features (FP32 master, updated by Adam)
  uint32_t* table_bf;  // kGuideLevels * T entries: both features as BF16 in one word (read by the encoder)
  float* table_grad;
  float* table_m;
  float* table_v;
  uint32_t log2T;
  int32_t res[kGuideLevels];  // grid resolution per level
  float* mlp;          // kGuideParams FP32 master: W_l [in][out] row-major, then biases
  float* mlp_grad;
  float* mlp_m;
  float* mlp_v;
  uint16_t* wf;        // BF16 W_l^T [out][in] per layer (forward operand), offsets guide_w_off
  uint16_t* wb;        // BF16 W_l [in][out] per layer (backward operand)
};

// Training batch (SoA floats): unit-cube position xyz, incoming direction xyz,
// sampled direction xyz, weight r = Li * f / pdf; n = *count (<= cap).
struct GuideBatch {
  float* s;            // 10 x cap floats
  uint32_t* count;
  uint32_t cap;
};
struct GuideTrain {
  uint16_t* act;       // 5 layers x cap x 64 BF16 layer inputs
  uint16_t* dy;        // cap x 64 BF16 output gradient of the current layer
  uint16_t* dy2;       // cap x 64 BF16 (ping-pong)
  float* denc;         // cap x 32 encoding gradient
  float* partial;      // waves x kGuideParams weight/bias gradient partials
  float* loss;         // [0] summed loss, [1] summed weight
  int32_t waves;
};

// Guided wavefront: per-path training record (one per slot and path buffer).
// state 0: none; 1: a guided vertex was sampled, the radiance baseline is
// taken when the next segment is shaded; 2: recording until the path ends.
struct WfGuideRec {
  uint32_t* state;
  float *px, *py, *pz;  // unit-cube position of the recorded vertex
  float *dx, *dy, *dz;  // incoming direction
  float *wx, *wy, *wz;  // sampled direction
  float* ratio;         // phase / mixture pdf of the sampled direction
  float *bx, *by, *bz;  // throughput after the vertex
  float *lx, *ly, *lz;  // radiance gathered when the next segment starts
};
constexpr int kGuideRecF = 16, kGuideRecU = 1;
struct WfGuide {
  GuideNet net;
  GuideBatch batch;
  WfGuideRec rec[2];
  int32_t train;      // collect training samples
  int32_t max_depth;  // scattering vertices with depth < max_depth use the guide
  float prob;         // probability of sampling the proposal at a guided vertex
  float* mix;         // 40 mixture floats per path slot (this iteration's guided vertices)
  GuideCache cache;   // cache.mix != nullptr: vertices read the tabulated proposal (no per-vertex network)
};

struct WfGuideHook {
  WfGuide g;
  GuideTrain train;
  float* cache_mix;     // writable cache table (re-tabulated after every training step), or nullptr
  float lr_table, lr_mlp, beta1, beta2, eps;
  uint32_t min_batch;   // a training step runs once this many samples are queued
  uint64_t steps;       // Adam steps (updated by the launcher)
  uint64_t samples;     // samples consumed
};
size_t wf_guide_bytes(uint32_t P);
void wf_guide_carve(void* base, uint32_t P, WfGuide& g);

hipError_t launch_guide_init(const GuideNet& g, uint64_t seed, hipStream_t st);
hipError_t launch_guide_pack(const GuideNet& g, hipStream_t st);  // FP32 master -> BF16 operands
// the proposal tabulated over the cache's cells and direction bins (into mix)
hipError_t launch_guide_bake(const GuideNet& g, const GuideCache& c, float* mix, hipStream_t st);
// outputs: n x 40 mixture parameters (weights, unit means, kappa)
hipError_t launch_guide_eval(const GuideNet& g, int n, const float* in6, float* out40, hipStream_t st);
// raw network outputs (n x 40) before the head, for parity tests
hipError_t launch_guide_raw(const GuideNet& g, int n, const float* in6, float* out40, hipStream_t st);
// one step on the batch: forward, loss gradient, backward into mlp_grad / table_grad (no update)
hipError_t launch_guide_grad(const GuideNet& g, const GuideBatch& b, const GuideTrain& t, hipStream_t st);
hipError_t launch_guide_adam(const GuideNet& g, float lr_table, float lr_mlp, float beta1, float beta2, float eps,
                             int step, hipStream_t st);

hipError_t launch_reduce_partials(const RenderArgs& a, double* film, hipStream_t st);
hipError_t launch_clear_rect(double* film, int W, int x0, int y0, int x1, int y1, hipStream_t st);
hipError_t launch_tonemap(const double* film, int npix, float inv_spp, float* pixels, hipStream_t st);
// Post-processes the film pixels of [x0, x0 + rw) x [y0, y0 + rh) into a packed rw x rh x 3 array.
hipError_t launch_tonemap_rect(const double* film, int W, int x0, int y0, int rw, int rh, float inv_spp, float* out,
                               hipStream_t st);

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
// row-major density -> bricked layout (grid_offset), zero padding
hipError_t launch_brick_grid(const float* lin, float* out, const int res[3], hipStream_t st);
hipError_t launch_majorants(const DevScene& s, float* maj, hipStream_t st);

}  // namespace ne
