// batch.hip -- persistent batched renderer: many render jobs of uneven size
// (own scene, camera, image size, spp) in one kernel launch.
//
// Work is cut into items: one item = a tile of tile_w x tile_h pixels of one
// job x samples_per_item consecutive samples.  Items of all jobs form one
// global queue (job by job; within a job, passes of up to kBatchPassSamples
// samples per pixel, tiles, sample chunks of the pass: batch_item_place); a wave takes the
// next item with one global atomic and hands its paths to free lanes.  A lane
// runs one bounce per loop iteration (path_bounce) and takes a new path as soon
// as its path ends, so neither uneven path lengths nor uneven jobs leave lanes
// idle until the queue is empty.  Every finished path adds its radiance to its
// job's 64-bit fixed-point film: the sums are independent of scheduling, so a
// job rendered in a batch is bitwise equal to the same job rendered alone.  The
// finishing lane also raises its item's completion timestamp, from which the
// per-job completion times are reduced after the launch.
#include "ne_launch.h"
#include "ne_trace.h"

namespace ne {

constexpr int kBatchBlock = 256;

__global__ void k_batch_init(BatchArgs b, size_t film_words) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t k = i; k < film_words; k += stride) b.films[k] = 0;
  for (size_t k = i; k < b.total_items; k += stride) b.item_time[k] = 0;
  if (i == 0) {
    *b.next_item = 0;
    *b.start_time = (unsigned long long)wall_clock64();
  }
}

// Item order within a job.  Tiles outer with every sample chunk of a tile in
// a row keeps the paths in flight on a few neighbouring tiles, which share
// voxels in the caches; at high spp those few tiles also take every film
// atomic in flight.  So the samples are cut into passes of kBatchPassSamples
// per pixel: pass outer, then tile, then the pass's chunks.
constexpr uint32_t kBatchPassSamples = 256;
__device__ __forceinline__ void batch_item_place(const BatchJob& J, uint32_t spi, unsigned long long local,
                                                 uint32_t& chunk, uint32_t& tile) {
  const uint32_t ntiles = J.tiles_x * J.tiles_y;
  const uint32_t cg = max(1u, kBatchPassSamples / spi);  // chunks per full pass
  const uint32_t full = J.chunks / cg, rem = J.chunks % cg;
  const unsigned long long per_pass = (unsigned long long)ntiles * cg;
  if (local < per_pass * full) {
    const uint32_t g = uint32_t(local / per_pass), r = uint32_t(local % per_pass);
    tile = r / cg;
    chunk = g * cg + r % cg;
  } else {  // the last, shorter pass
    const uint32_t r = uint32_t(local - per_pass * full);
    tile = r / rem;
    chunk = full * cg + r % rem;
  }
}

// Job of global item q: the last job whose first item is <= q.
__device__ __forceinline__ int batch_job_of(const unsigned long long* first, int njobs, unsigned long long q) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= q) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// At least 3 waves per SIMD: the register budget drops from 185 to 168
// VGPRs with 8 dwords spilled; measured on MI355X, batch48 ran 1636 ms per
// launch at 2 waves and 1368 ms at 3 (4 waves spill 48 dwords: 1550 ms).
// SURF = false: every job is a volume lit by the sun and the environment
// (no surfaces, no emitters), and the bounce carries no surface code.
// GUIDED: jobs with a guide sample its tabulated proposal at medium vertices.
template <bool SURF, bool GUIDED>
__global__ __launch_bounds__(kBatchBlock, 3) void k_batch(BatchArgs b) {
  __shared__ uint32_t lds_stack[kShortStack * kBatchBlock];
  __shared__ unsigned long long first[kMaxBatchJobs];
  for (int j = threadIdx.x; j < b.njobs; j += blockDim.x) first[j] = b.jobs[j].first_item;
  __syncthreads();
  const int lane = __lane_id();
  uint32_t* lds = lds_stack + threadIdx.x;
  Count<false> cnt;
  // current item of the wave (identical in every lane still looping)
  int ij = 0;
  uint32_t ix0 = 0, iy0 = 0, iw = 0, ih = 0, is0 = 0, item_n = 0, item_next = 0;
  unsigned long long iq = 0;
  bool empty = false;
  // the lane's path
  bool have = false;
  PathState<float> ps;
  PathRng rng;
  int lj = 0;
  uint32_t lpix = 0;
  unsigned long long lq = 0;
  while (true) {
    // hand paths of the wave's item to lanes without one; fetch items as needed
    bool need = !have;
    while (true) {
      const unsigned long long m = __ballot(need);
      if (m == 0 || empty) break;
      if (item_next >= item_n) {
        const int leader = __ffsll((long long)__ballot(true)) - 1;
        unsigned long long q = 0;
        if (lane == leader) q = atomicAdd(b.next_item, 1ull);
        q = __shfl(q, leader);
        if (q >= b.total_items) {
          empty = true;
          break;
        }
        ij = batch_job_of(first, b.njobs, q);
        const BatchJob& J = b.jobs[ij];
        const unsigned long long local = q - first[ij];
        uint32_t chunk, tile;
        batch_item_place(J, b.spi, local, chunk, tile);
        ix0 = (tile % J.tiles_x) * b.tile_w;
        iy0 = (tile / J.tiles_x) * b.tile_h;
        iw = min(b.tile_w, uint32_t(J.W) - ix0);
        ih = min(b.tile_h, uint32_t(J.H) - iy0);
        is0 = chunk * b.spi;
        item_n = iw * ih * min(b.spi, uint32_t(J.spp) - is0);
        item_next = 0;
        iq = q;
        continue;
      }
      const uint32_t k = item_next + uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
      item_next = min(item_n, item_next + uint32_t(__popcll(m)));
      if (need && k < item_n) {
        const BatchJob& J = b.jobs[ij];
        const uint32_t np = iw * ih;
        const uint32_t px = ix0 + (k % np) % iw, py = iy0 + (k % np) / iw;
        const uint32_t smp = is0 + k / np;
        lj = ij;
        lq = iq;
        lpix = py * uint32_t(J.W) + px;
        rng.init(J.seed, lpix, smp, J.job);
        const float ju = rng.next(), jv = rng.next();
        const float u = (float(px) + ju) / float(J.W), v = (float(py) + jv) / float(J.H);
        const Vec3<float> pos = v3f<float>(J.cam.pos), ll = v3f<float>(J.cam.lower_left);
        const Vec3<float> hz = v3f<float>(J.cam.horizontal), vt = v3f<float>(J.cam.vertical);
        path_init(ps, pos, -normalize(ll + u * hz + v * vt - pos));
        have = true;
        need = false;
      }
    }
    if (!have) break;  // queue empty and no path left for this lane
    const BatchJob& J = b.jobs[lj];
    if (path_bounce<float, false, SURF, GUIDED>(J.scene, ps, J.max_bounces, rng, lds, kBatchBlock, cnt, &J.guide)) {
      film_add(J.film, lpix, ps.L.x, ps.L.y, ps.L.z);
      atomicMax(&b.item_time[lq], (unsigned long long)wall_clock64());
      have = false;
    }
  }
}

// Per-job completion time: latest item timestamp of the job minus the launch
// start, in milliseconds (one block per job).
__global__ void k_batch_job_times(BatchArgs b, double ticks_per_ms) {
  __shared__ unsigned long long red[256];
  const BatchJob& J = b.jobs[blockIdx.x];
  const unsigned long long n = (unsigned long long)J.tiles_x * J.tiles_y * J.chunks;
  unsigned long long m = 0;
  for (unsigned long long k = threadIdx.x; k < n; k += blockDim.x) {
    const unsigned long long t = b.item_time[J.first_item + k];
    m = t > m ? t : m;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < unsigned(o)) red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + o] ? red[threadIdx.x] : red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const unsigned long long t0 = *b.start_time;
    b.job_ms[blockIdx.x] = red[0] > t0 ? float(double(red[0] - t0) / ticks_per_ms) : 0.0f;
  }
}

// fixed-point film -> double sums
__global__ void k_fixed_to_double(const long long* in, double* out, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = double(in[i]) * (1.0 / 1073741824.0);
}

static const void* batch_kernel(bool surfaces, bool guided) {
  if (surfaces) return guided ? reinterpret_cast<const void*>(&k_batch<true, true>)
                              : reinterpret_cast<const void*>(&k_batch<true, false>);
  return guided ? reinterpret_cast<const void*>(&k_batch<false, true>)
                : reinterpret_cast<const void*>(&k_batch<false, false>);
}

hipError_t batch_blocks_per_cu(bool surfaces, bool guided, int* blocks) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, batch_kernel(surfaces, guided), kBatchBlock, 0);
}

hipError_t launch_batch(const BatchArgs& b, bool surfaces, bool guided, size_t film_words, uint32_t grid_blocks,
                        double ticks_per_ms, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_init, dim3(2048), dim3(256), 0, st, b, film_words);
  const dim3 grid(grid_blocks), blk(kBatchBlock);
  if (surfaces && guided) hipLaunchKernelGGL((k_batch<true, true>), grid, blk, 0, st, b);
  else if (surfaces) hipLaunchKernelGGL((k_batch<true, false>), grid, blk, 0, st, b);
  else if (guided) hipLaunchKernelGGL((k_batch<false, true>), grid, blk, 0, st, b);
  else hipLaunchKernelGGL((k_batch<false, false>), grid, blk, 0, st, b);
  hipLaunchKernelGGL(k_batch_job_times, dim3(uint32_t(b.njobs)), dim3(256), 0, st, b, ticks_per_ms);
  return hipGetLastError();
}

hipError_t launch_fixed_to_double(const long long* in, double* out, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_fixed_to_double, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, st, in, out, n);
  return hipGetLastError();
}

}  // namespace ne
