/*
 * ne_gpu.h -- C ABI of the MI355X (gfx950) path-tracing backend for
 * NarvalEngine.  Plain C types only: pointers, sizes, POD structs.
 *
 * The library replaces the CPU hot path
 *   OfflineEngine::renderTile  (src/core/OfflineEngine.cpp:54-76)
 *     -> Camera::getRayPassingThrough (src/core/Camera.cpp:140-144)
 *     -> VolumetricPathIntegrator::Li (src/integrators/VolumetricPathIntegrator.cpp:176-301)
 * behind the reference's Integrator / tile-render interfaces
 * (src/integrators/Integrator.h:13-25, src/core/OfflineEngine.h:35-59).
 *
 * Ownership: the caller owns every host buffer passed in or out; the library
 * owns device memory behind the opaque handles; descriptors are copied at
 * upload.  Errors: every entry point returns an ne_status; nothing aborts or
 * throws across the ABI (the reference LOG(FATAL)s, e.g.
 * src/io/SceneReader.cpp:24-41, src/materials/GridMedia.cpp:20).
 * Threading: calls on one context are serialised by an internal mutex and
 * run on the context's single HIP stream; distinct contexts are independent.
 */
#ifndef NE_GPU_H
#define NE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t ne_status;
enum {
  NE_OK = 0,
  NE_ERR_INVALID_ARG = 1,
  NE_ERR_HIP = 2,
  NE_ERR_OUT_OF_MEMORY = 3,
  NE_ERR_NO_DEVICE = 4,
  NE_ERR_UNSUPPORTED = 5,
  NE_ERR_PARSE = 6,
  NE_ERR_IO = 7
};

typedef struct ne_ctx ne_ctx;
typedef struct ne_scene ne_scene;
typedef struct ne_film ne_film;

/* Material record: 12 floats
 *   [0] type (NE_MAT_*), [1..3] albedo, [4] roughness, [5] metallic,
 *   [6..8] emitted radiance, [9..11] unused.
 * Glossy = GlossyBSDF (src/core/GlossyBSDF.cpp:4-48, alpha = roughness^2);
 * emitter = DiffuseLight (src/lights/DiffuseLight.cpp:4-27). */
enum { NE_MAT_GLOSSY = 0, NE_MAT_EMITTER = 1 };
/* Environment: none, constant radiance env_c0, a sky gradient (zenith
 * env_c0, horizon env_c1, below-horizon env_c2), or an InfiniteAreaLight
 * (src/lights/InfiniteAreaLight.h:8-130) latitude-longitude map: env_w x
 * env_h RGB floats, row-major, x fastest, row 0 at theta = 0 of the map's
 * z-up frame, importance-sampled. */
enum { NE_ENV_NONE = 0, NE_ENV_CONSTANT = 1, NE_ENV_SKY = 2, NE_ENV_MAP = 3 };

/* Flattened world-space scene.  Mirrors what SceneReader builds
 * (src/io/SceneReader.cpp:224-648) after InstancedModel transforms:
 *   triangles      9 floats each (v0, v1, v2)         Triangle (src/primitives/Triangle.cpp)
 *   spheres        4 floats each (centre, radius)     Sphere (src/primitives/Sphere.cpp)
 *   rectangles    12 floats each (corner p0, edge eu, edge ev, normal n);
 *                 the reference's unit quad [-.5,.5]^2 at z=0 with normal
 *                 (0,0,-1) under its instance transform (SceneReader.cpp:407-419)
 *   volume         one GridMedia (src/materials/GridMedia.cpp) on the unit
 *                 cube [-.5,.5]^3 under vol_xf (row-major 4x4, OCS->WCS);
 *                 grid is x-fastest R32F (to1D, src/utils/Math.h:104-106);
 *                 with vol_homogeneous set, a HomogeneousMedia
 *                 (src/materials/HomogeneousMedia.cpp:5-51) of density
 *                 vol_mult on the same cube instead (grid and vol_res unused)
 *   sun            DirectionalLight (src/lights/DirectionalLight.cpp), sun_dir is
 *                 the direction light travels, sun_E its radiance scale
 * The same layout is accepted by the test oracle, so one host-side
 * descriptor drives both. */
typedef struct ne_scene_desc {
  const float* tri_v;
  const int32_t* tri_mat;
  uint32_t n_tri;
  const float* sph;
  const int32_t* sph_mat;
  uint32_t n_sph;
  const float* rect;
  const int32_t* rect_mat;
  uint32_t n_rect;
  const float* mat;
  uint32_t n_mat;
  int32_t has_volume;
  float vol_xf[16];
  float vol_inv[16];
  int32_t vol_res[3];
  const float* grid;
  float vol_mult;
  float vol_sigma_s[3];
  float vol_sigma_a[3];
  float vol_g;
  int32_t maj_res; /* majorant cells per axis; 1 = the reference's single global majorant */
  int32_t has_sun;
  float sun_dir[3];
  float sun_E[3];
  int32_t env_type;
  float env_c0[3];
  float env_c1[3];
  float env_c2[3];
  int32_t vol_homogeneous; /* 1: constant density vol_mult inside the cube; analytic transmittance */
  const float* env_map;    /* NE_ENV_MAP: env_w * env_h * 3 floats */
  int32_t env_w;
  int32_t env_h;
  const float* point;       /* point lights: 3 floats each (position), Point (src/primitives/Point.cpp) */
  const int32_t* point_mat; /* their emitter materials (radiance = intensity) */
  uint32_t n_point;
  int32_t maj_no_skip; /* 1: the majorant DDA reads every cell it crosses; 0: empty cells carry their
                          L1 distance to the nearest non-empty cell and runs of them are stepped
                          over without reads (same image) */
} ne_scene_desc;

/* Pinhole camera: Camera(lookFrom, lookAt, up, vfov, aspect, ...)
 * (src/core/Camera.cpp:7-26); ray through (u,v) per getRayPassingThrough
 * (src/core/Camera.cpp:140-144) with the lens offset at zero. */
typedef struct ne_camera_desc {
  float position[3];
  float look_at[3];
  float up[3];
  float vfov;
  float aspect;
} ne_camera_desc;

enum { NE_FP32 = 0, NE_FP64 = 1 };
enum {
  NE_FLAG_COUNT = 1,      /* collect event counters (separate, slower pass) */
  NE_FLAG_TIME = 2,       /* time the path-tracing kernels with HIP events on the context stream */
  NE_FLAG_STAGE_TIME = 4, /* also time every kernel launch and sum the times per stage (NE_STAGE_*) */
  NE_FLAG_SORT = 8,       /* wavefront: shade each chunk's paths ordered by (event kind, material) */
  NE_FLAG_FIXED_FILM = 16, /* wavefront: accumulate through 2^-30 fixed-point film atomics, the path taken
                             anyway when the render's per-sample radiance slots (12 B each) exceed 16 GiB */
  NE_FLAG_SPATIAL = 32,    /* wavefront: the delta- and ratio-tracking stages take each chunk's rays in
                              Morton order of their origins (same paths, same image) */
  NE_FLAG_WIDE_OFFSETS = 64, /* wavefront: read the density grid by 64-bit offsets even under 4 GiB, the
                                kernels a grid of 4 GiB or more takes (same addresses, same image) */
  NE_FLAG_TEST_PLAN_STALL = 128 /* test hook for the failure path: the wavefront plan's look-back waits for
                                   words that never come, its bounded wait gives up after 1024 spins, and
                                   the render fails with NE_ERR_HIP (a launch time-out) instead of hanging.
                                   Only pools of more than 2^18 slots (1024 chunks) have a look-back:
                                   smaller pools and the megakernel render as usual */
};

/* Kernel stages reported by ne_get_stage_times. */
enum {
  NE_STAGE_MEGAKERNEL = 0, /* one thread per (pixel, sample chunk), whole paths */
  NE_STAGE_PLAN = 1,       /* wavefront: hand out new path ids per chunk (until every path is issued) */
  NE_STAGE_REGEN = 2,      /* wavefront: camera rays into free slots (volume-only scenes: inside MEDIUM) */
  NE_STAGE_TRACE = 3,      /* wavefront: BVH closest hit */
  NE_STAGE_MEDIUM = 4,     /* wavefront: delta tracking */
  NE_STAGE_SHADE = 5,      /* wavefront: emission, NEE sample, scattering, compaction (volume-only
                              scenes without training or counting: also the shadow rays' ratio tracking) */
  NE_STAGE_SHADOW = 6,     /* wavefront: occlusion + ratio tracking (when not in SHADE) */
  NE_STAGE_OTHER = 7,      /* initialisation, film reduction / flush */
  NE_STAGE_GUIDE = 8,      /* guided wavefront: learned proposal evaluation (MFMA) */
  NE_STAGE_TRAIN = 9,      /* guided wavefront: online training steps */
  NE_STAGE_COUNT = 10
};

/* One render call: samples [sample_offset, sample_offset + spp) of every pixel
 * in [x0,x1) x [y0,y1) of a width x height image (SceneSettings,
 * src/core/Settings.h:9-14).  Path length: max_bounces scatter vertices. */
typedef struct ne_render_params {
  int32_t width, height;
  int32_t x0, y0, x1, y1;
  int32_t spp;
  int32_t max_bounces;
  uint64_t seed;
  uint32_t sample_offset;
  uint32_t job;
  int32_t precision; /* NE_FP32 (product) or NE_FP64 (converged-oracle mode) */
  int32_t flags;
  int32_t samples_per_thread; /* 0 = library default */
  int32_t block_size;         /* 0 = library default */
  int32_t occupancy;          /* megakernel FP32 variant built for >= this many waves per SIMD: 0, 4, 6 */
  int32_t mode;               /* NE_MODE_*: 0 = auto (wavefront for FP32, megakernel for FP64) */
  int32_t pool_size;          /* wavefront path slots (rounded up to 256); 0 = library default
                                 (2^23 for surface-only scenes and for renders of 2^28 paths
                                 or more, else 2^22) */
} ne_render_params;

enum { NE_MODE_AUTO = 0, NE_MODE_MEGAKERNEL = 1, NE_MODE_WAVEFRONT = 2 };

typedef struct ne_counters {
  uint64_t paths;
  uint64_t ext_rays;        /* closest-hit queries */
  uint64_t shadow_rays;     /* transmittance queries */
  uint64_t density_lookups; /* trilinear grid lookups (32 B each) */
  uint64_t majorant_cells;  /* majorant-grid cells visited (4 B each) */
  uint64_t bvh_nodes;       /* BVH nodes fetched (64 B each) */
  uint64_t prim_tests;      /* primitive intersection tests */
  /* the part of the four counts above spent on shadow rays (wavefront
   * pipeline; zero for the megakernel, which does not split them) */
  uint64_t shadow_density_lookups;
  uint64_t shadow_majorant_cells;
  uint64_t shadow_bvh_nodes;
  uint64_t shadow_prim_tests;
  uint64_t guide_queries;   /* learned-proposal evaluations (guided rendering; 34,176 MLP FLOP each) */
} ne_counters;

typedef struct ne_scene_info {
  uint32_t bvh_nodes;
  uint32_t bvh_depth;
  uint32_t bvh_leaf_prims_max;
  uint32_t maj_res;
  uint64_t device_bytes;
  float grid_max_density;
  float pad;
} ne_scene_info;

/* ---- context ------------------------------------------------------------ */
ne_status ne_ctx_create(int32_t device, ne_ctx** out);
/* Releases the caller's context handle.  Scenes and films created on the
 * context keep it alive until they are released, in any order. */
void ne_ctx_destroy(ne_ctx* ctx);
/* Last error message of this context (static string if ctx is NULL). */
const char* ne_last_error(const ne_ctx* ctx);
ne_status ne_synchronize(ne_ctx* ctx);
/* Size of ne_scene_desc as compiled into the library (ABI check). */
int32_t ne_sizeof_scene_desc(void);

/* ---- scene ---------------------------------------------------------------
 * Replaces SceneReader's Scene/BVH/GridMedia construction
 * (src/io/SceneReader.cpp:10-54, src/primitives/BVH.cpp:6-106,
 * src/materials/GridMedia.cpp:4-13): builds a binned-SAH BVH and the
 * majorant grid on the host and uploads everything to HBM. */
ne_status ne_scene_upload(ne_ctx* ctx, const ne_scene_desc* desc, ne_scene** out);
void ne_scene_release(ne_scene* scene);
ne_status ne_scene_get_info(const ne_scene* scene, ne_scene_info* out);

/* ---- film (device FP64 accumulator, x-fastest, 3 doubles per pixel) ------ */
ne_status ne_film_create(ne_ctx* ctx, int32_t width, int32_t height, ne_film** out);
void ne_film_release(ne_film* film);
ne_status ne_film_clear(ne_film* film);
/* Copies the per-pixel radiance sums (width*height*3 doubles) to the host. */
ne_status ne_film_read(ne_film* film, double* sum_rgb);
/* OfflineEngine::postProcessing (src/core/OfflineEngine.cpp:27-52) of
 * sum/spp into pixels (width*height*3 floats, row-major, x fastest). */
ne_status ne_film_tonemap(ne_film* film, int32_t spp, float* pixels);

/* ---- render --------------------------------------------------------------
 * Enqueues the render on the context stream; returns when enqueued unless an
 * argument error is found.  ne_synchronize waits for completion. */
ne_status ne_render(ne_ctx* ctx, const ne_scene* scene, const ne_camera_desc* cam, const ne_render_params* params,
                    ne_film* film);
/* Tile call with the reference's tiling: 40 x 10 tiles of (W/40) x (H/10)
 * pixels, tile `index` = (index % 40, index / 40) (OfflineEngine.cpp:18-21,
 * 54-58).  Clears the tile's film sums, renders params->spp samples of that
 * tile, synchronises, and writes post-processed pixels of the tile into
 * pixels (W*H*3 floats); other pixels of the array are left untouched.
 * Below 40 x 10 pixels the tile size is 0 on an axis: as in renderTile, the
 * tile is empty and the call renders and writes nothing (NE_OK).  Indices
 * outside [0, 400) and spp <= 0 are argument errors. */
ne_status ne_render_tile(ne_ctx* ctx, const ne_scene* scene, const ne_camera_desc* cam,
                         const ne_render_params* params, int32_t index, ne_film* film, float* pixels);
/* ne_render without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It runs the
 * megakernel pipeline only (params->mode AUTO or MEGAKERNEL, FP32 or FP64;
 * no NE_FLAG_COUNT / NE_FLAG_STAGE_TIME) and gives the same film sums as
 * ne_render with NE_MODE_MEGAKERNEL and the same params.  When the samples
 * of a pixel are split into chunks (spp > samples per thread), the chunk
 * partial sums, (x1-x0)*(y1-y0)*chunks*3 doubles, live in a static device
 * array of 32 MiB per GPU, held until the process ends; larger requests are
 * argument errors.  Every call on a GPU shares that array: calls that overlap
 * in time on one GPU (on different contexts) race. */
ne_status ne_render_noalloc(ne_ctx* ctx, const ne_scene* scene, const ne_camera_desc* cam,
                            const ne_render_params* params, ne_film* film);
/* Duration in milliseconds of the path-tracing kernel of the last render
 * issued with NE_FLAG_TIME (waits for it to finish). */
ne_status ne_get_kernel_time(ne_ctx* ctx, float* ms);
/* Per-stage kernel time of the last render issued with NE_FLAG_STAGE_TIME
 * (waits for it to finish): ms[NE_STAGE_COUNT] summed durations (each launch
 * timed from the event before it to the event after it on the context
 * stream) and launches[NE_STAGE_COUNT] launch counts. */
ne_status ne_get_stage_times(ne_ctx* ctx, float* ms, uint32_t* launches);
/* Event counters of the last render issued with NE_FLAG_COUNT. */
ne_status ne_get_counters(ne_ctx* ctx, ne_counters* out);
/* Integrator::Li (src/integrators/Integrator.h:23) for a batch of n rays
 * (origin xyz, direction xyz): radiance of one path per ray, path stream i
 * keyed by (seed, pixel = i, sample = sample_index). */
ne_status ne_li(ne_ctx* ctx, const ne_scene* scene, int32_t n, const float* rays, int32_t max_bounces, uint64_t seed,
                uint32_t sample_index, float* radiance);

/* ---- batched rendering ----------------------------------------------------
 * Many render jobs of uneven size (own scene, camera, image size, spp) in one
 * call.  The reference renders one scene per OfflineEngine
 * (src/core/OfflineEngine.cpp:8-25, 78-145); this entry point has no
 * counterpart there.  Two executors:
 *   NE_BATCH_PERSISTENT: one persistent kernel launch, captured in a hipGraph;
 *     work items (a tile of one job x a range of samples) go through one
 *     global queue and each lane steps its path one bounce at a time.
 *   NE_BATCH_WAVEFRONT: jobs, longest first, on NE_BATCH_LANES streams, each
 *     with its own path pool, through the wavefront pipeline (the FP32
 *     ne_render path); one job's tail overlaps the others' work.
 * Per-job sums are 64-bit fixed point (2^-30 units), so a job's result is
 * bitwise independent of the batch it runs in, of the item shape and of the
 * order of the lanes (for a given executor).  Scenes must outlive the batch. */
typedef struct ne_job {
  const ne_scene* scene;
  ne_camera_desc camera;
  int32_t width, height;
  int32_t spp;
  int32_t max_bounces;
  uint64_t seed;
  uint32_t job; /* RNG stream word: paths are keyed (seed; pixel, sample, job) */
  uint32_t reserved;
  /* NULL: plain phase sampling.  Else a guide with a tabulated proposal
   * (cache_res > 0) from the same context: the job's medium vertices sample it
   * as ne_render_guided does (guided+batched; no training inside the batch). */
  const struct ne_guide* guide;
} ne_job;
/* Executor and, for NE_BATCH_PERSISTENT, the work-item shape: tile_w x tile_h
 * pixels x samples_per_item samples (ignored by NE_BATCH_WAVEFRONT). */
enum { NE_BATCH_PERSISTENT = 0, NE_BATCH_WAVEFRONT = 1 };
#define NE_BATCH_LANES 4 /* streams of the wavefront executor */
typedef struct ne_batch_config {
  int32_t tile_w, tile_h, samples_per_item;
  int32_t executor; /* NE_BATCH_PERSISTENT or NE_BATCH_WAVEFRONT */
} ne_batch_config;
typedef struct ne_batch ne_batch;
ne_status ne_batch_create(ne_ctx* ctx, const ne_job* jobs, int32_t n, ne_batch** out);
void ne_batch_release(ne_batch* batch);
/* Times each candidate (8 item shapes of the persistent kernel and the
 * wavefront executor) on a calibration run (every job at ceil(spp /
 * spp_divisor) samples); with spp_divisor > 1 the fastest persistent shape and
 * the wavefront executor are then compared on one full-spp run each.  Keeps
 * the winner for later runs and returns it with its calibration time. */
ne_status ne_batch_autotune(ne_batch* batch, int32_t spp_divisor, ne_batch_config* best, float* best_ms);
/* Renders every job at full spp (cfg NULL: the tuned configuration, else the
 * persistent kernel with 8 x 8 x 4 items) and waits; ms = the graph launch's
 * duration (HIP events) or, for the wavefront executor, the wall time from the
 * first job's start to the last one's end. */
ne_status ne_batch_run(ne_batch* batch, const ne_batch_config* cfg, float* ms);
/* Completion time of each job in the last run, ms from the run's start. */
ne_status ne_batch_job_times(ne_batch* batch, float* ms);
/* Linear radiance sums of job `job` (width * height * 3 doubles). */
ne_status ne_batch_read(ne_batch* batch, int32_t job, double* sum_rgb);

/* ---- learned directional guide --------------------------------------------
 * A per-position, per-direction proposal for medium scattering vertices: the
 * unit-cube position of the volume through a multiresolution hash grid
 * (16 levels x 2 features, 2^log2_table entries per level) plus the incoming
 * direction feed an MLP 35 -> 64 x 4 (ReLU) -> 40 on BF16 MFMA tiles (FP32
 * accumulation) that outputs an 8-lobe von Mises-Fisher mixture.  It is trained
 * online with Adam on the tracer's own radiance samples and combined with HG
 * sampling by one-sample MIS ((1 - guide_prob) HG + guide_prob guide), so the estimator stays unbiased.
 * New component: the reference has no learned sampling. */
typedef struct ne_guide_config {
  int32_t log2_table; /* hash entries per level = 2^log2_table (default 19, range 10..22) */
  int32_t base_res;   /* coarsest level resolution (default 16) */
  int32_t max_res;    /* finest level resolution (default 512) */
  int32_t batch;      /* training samples per step (default 65536) */
  float lr_table, lr_mlp, beta1, beta2, eps; /* Adam (defaults 1e-2, 1e-3, 0.9, 0.99, 1e-8) */
  int32_t guided_vertices; /* the first this many scattering vertices of a path use the guide
                              (default 1): each mixture-sampled vertex scales the throughput by
                              phase / mixture pdf <= 1 / (1 - guide_prob), so the product stays
                              bounded */
  float guide_prob;        /* probability of sampling the guide at a guided vertex (default 0.25,
                              range [0, 0.9]); the rest samples HG (one-sample MIS) */
  uint64_t seed;      /* parameter initialisation */
  int32_t cache_res;  /* > 0: the proposal is tabulated on cache_res^3 cells of the volume's unit cube x
                         cache_dirs^2 octahedral bins of the incoming direction (re-tabulated after every
                         training step) and vertices read their entry instead of evaluating the network;
                         needed by the guided megakernel and guided batch jobs.  0 (default): per-vertex
                         network evaluation in the wavefront.  Range [0, 128]. */
  int32_t cache_dirs; /* direction bins per axis (default 4, range [1, 16]) */
} ne_guide_config;
typedef struct ne_guide_stats {
  uint64_t steps;     /* Adam steps taken */
  uint64_t samples;   /* training samples consumed */
  float last_loss;    /* weighted loss of the last step */
  float pad;
} ne_guide_stats;
typedef struct ne_guide ne_guide;
/* cfg NULL: defaults.  Parameters are initialised from cfg->seed. */
ne_status ne_guide_create(ne_ctx* ctx, const ne_guide_config* cfg, ne_guide** out);
void ne_guide_release(ne_guide* guide);
/* Re-initialises the parameters and the optimiser state. */
ne_status ne_guide_reset(ne_guide* guide);
ne_status ne_guide_get_stats(ne_guide* guide, ne_guide_stats* out);
/* n queries (unit-cube position xyz, direction xyz) -> 40 floats each: per
 * lobe (weight, unit mean xyz, kappa); raw != 0: the network's 40 raw outputs. */
ne_status ne_guide_eval(ne_guide* guide, int32_t n, const float* queries, int32_t raw, float* out);
/* One training step on n <= batch samples of 10 floats (position xyz,
 * incoming direction xyz, sampled direction xyz, weight r >= 0): forward,
 * gradient of -sum r log q / sum r and, if update != 0, an Adam step.  loss:
 * the weighted loss before the step. */
ne_status ne_guide_train_step(ne_guide* guide, int32_t n, const float* samples, int32_t update, float* loss);
/* The tabulated proposal (cache_res > 0), brought up to date with the
 * parameters first: entries = cache_res^3 * cache_dirs^2 records of 40 floats
 * (per lobe: weight, unit mean xyz, kappa), cell (x fastest) outer, direction
 * bin inner.  Cell (i, j, k) is evaluated at the unit-cube point
 * ((i + .5) / cache_res, ...), bin b = by * cache_dirs + bx at the octahedral
 * point ((bx + .5) / cache_dirs, (by + .5) / cache_dirs) of [0, 1]^2, where a
 * unit direction d maps to (p + 1) / 2 with p = d.xy / |d|_1, folded as
 * (1 - |p.yx|) sign(p) for d.z < 0.  mix NULL: only the entry count. */
ne_status ne_guide_get_cache(ne_guide* guide, float* mix, uint64_t* entries);
/* ne_render with the guide: the first guided_vertices medium scattering
 * vertices of every path sample the one-sample MIS mixture
 * (1 - guide_prob) HG + guide_prob guide, later ones HG (FP32).  The
 * wavefront pipeline (mode AUTO or WAVEFRONT) evaluates the network per
 * vertex, or reads the tabulated proposal when cache_res > 0; the megakernel
 * (mode MEGAKERNEL) needs the tabulation and does not train.  With train != 0
 * (wavefront) one guided vertex of every path, chosen uniformly, is recorded;
 * when the path ends it becomes a training sample (weight m Li phase / pdf, m
 * the path's guided-vertex count), and an Adam step runs whenever half a batch
 * is queued (the tabulation is refreshed after each step). */
ne_status ne_render_guided(ne_ctx* ctx, const ne_scene* scene, const ne_camera_desc* cam,
                           const ne_render_params* params, ne_film* film, ne_guide* guide, int32_t train);
/* Sizes: hash-table floats, MLP floats (weights [in][out] per layer, padded to
 * 48/64, then biases) and the 16 level resolutions. */
ne_status ne_guide_sizes(ne_guide* guide, uint64_t* table_floats, int32_t* mlp_floats, int32_t* level_res);
ne_status ne_guide_get_params(ne_guide* guide, float* table, float* mlp);
ne_status ne_guide_set_params(ne_guide* guide, const float* table, const float* mlp);
/* Gradients left by the last train step taken with update = 0 (an update
 * consumes and clears them). */
ne_status ne_guide_get_grads(ne_guide* guide, float* table_grad, float* mlp_grad);

/* ---- per-function parity entry points (device code, host buffers) ------- */
enum {
  NE_TEST_AABB = 0,     /* AABB::intersect      src/primitives/AABB.cpp:48-79      prim: min xyz, max xyz   */
  NE_TEST_TRIANGLE = 1, /* Triangle::intersect  src/primitives/Triangle.cpp:49-81  prim: v0, v1, v2         */
  NE_TEST_SPHERE = 2,   /* Sphere::intersect    src/primitives/Sphere.cpp:21-44    prim: centre, radius     */
  NE_TEST_RECTANGLE = 3 /* Rectangle::intersect src/primitives/Rectangle.cpp:50-76 prim: vA, vB, normal     */
};
/* Per-ray primitive test in the reference's semantics; t0/t1 = tNear/tFar. */
ne_status ne_test_primitive(ne_ctx* ctx, int32_t kind, int32_t n, const float* rays, const float* prims, int32_t* hit,
                            float* t0, float* t1);
/* Closest surface hit in (tmin, tmax) through the device BVH; type -1 = miss. */
ne_status ne_test_closest(ne_ctx* ctx, const ne_scene* scene, int32_t n, const float* rays, float tmin, float tmax,
                          float* t, int32_t* type, int32_t* index);
/* Shadow transmittance (surface occlusion x ratio tracking,
 * GridMedia::Tr src/materials/GridMedia.cpp:45-69) along n rays to distance
 * tmax[i]; stream i keyed by (seed, pixel = i, sample 0). */
ne_status ne_test_transmittance(ne_ctx* ctx, const ne_scene* scene, int32_t n, const float* rays, const float* tmax,
                                uint64_t seed, int32_t precision, double* tr);
/* Delta tracking (GridMedia::sample, src/materials/GridMedia.cpp:71-98) along n
 * rays through the volume; collision flag and distance. */
ne_status ne_test_delta_track(ne_ctx* ctx, const ne_scene* scene, int32_t n, const float* rays, uint64_t seed,
                              int32_t precision, int32_t* scattered, double* t);
/* Trilinear density (GridMedia::interpolatedDensity, GridMedia.cpp:25-38) at
 * n grid-space points of the scene's volume (even points through the lookup
 * for grids under 4 GiB, odd points through the general one). */
ne_status ne_test_density(ne_ctx* ctx, const ne_scene* scene, int32_t n, const float* gcs, float* out);
/* The scene's majorant grid as the device holds it (maj_res^3 floats, x
 * fastest): positive per-cell majorants, and in empty cells 0 or, with empty-
 * space skipping, -D (D the L1 cell distance to the nearest non-empty cell).
 * No reference counterpart: GridMedia keeps one global majorant
 * (src/materials/GridMedia.cpp:82). */
ne_status ne_test_majorants(ne_ctx* ctx, const ne_scene* scene, float* out, int64_t capacity);
/* Phase / BSDF functions:
 *   NE_TEST_HG_EVAL:   in a[3n], b[3n], g[n]           -> out[n]   HG::eval (Medium.h:116-128)
 *   NE_TEST_HG_SAMPLE: in u[2n], g[n]                  -> out[3n]  HG::sample local (Medium.h:96-114)
 *   NE_TEST_GGX_D:     in h[3n], alpha[n]              -> out[n]   GGXDistribution::D (Microfacet.cpp:5-15)
 *   NE_TEST_GGX_PDF:   in wo[3n], h[3n], alpha[n]      -> out[n]   GGXDistribution::pdf (Microfacet.cpp:58-65)
 *   NE_TEST_GLOSSY:    in mat[5n], wo[3n], wi[3n], u[3n] -> out[7n] (f rgb, pdf, sampled wi)
 *                      corrected GGX+Lambert lobe used by the integrator
 *   NE_TEST_HG_FWD:    in d[3n], u[2n], g[n]           -> out[4n]  (sampled wi, pdf) forward-convention HG */
enum { NE_TEST_HG_EVAL = 0, NE_TEST_HG_SAMPLE = 1, NE_TEST_GGX_D = 2, NE_TEST_GGX_PDF = 3, NE_TEST_GLOSSY = 4,
       NE_TEST_HG_FWD = 5 };
ne_status ne_test_bsdf(ne_ctx* ctx, int32_t kind, int32_t n, const float* in0, const float* in1, const float* in2,
                       const float* in3, float* out);
/* Device Philox-4x32-10 uniform stream: n uniforms of stream (seed, pixel, sample, job). */
ne_status ne_test_rng(ne_ctx* ctx, uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t job, int32_t n,
                      float* out);

/* Host-only: builds the BVH that ne_scene_upload would build for desc and
 * returns it (no device needed).  nodes gets 16 words per 4-wide node in the
 * device layout: float origin[3]; uint32 exponents (byte a = e_a + 127);
 * uint32 qlo.x, qlo.y, qlo.z, qhi.x, qhi.y, qhi.z (one byte per child: child
 * c's box on axis a is [origin_a + qlo_a[c] 2^e_a, origin_a + qhi_a[c] 2^e_a]
 * in FP32); uint32 meta (byte c: 0 inner child, else 0x80 | type << 4 |
 * count); uint32 0; int32 child[4] (inner node index, leaf first slot, or -1
 * for an empty slot).  refs gets one (type << 30 | original primitive index)
 * per leaf slot; depth the number of wide-node levels.
 * Call with NULL nodes/refs to query n_nodes / n_refs first. */
ne_status ne_bvh_build(const ne_scene_desc* desc, float* nodes, uint32_t* refs, uint32_t* n_nodes, uint32_t* n_refs,
                       uint32_t* depth);

/* ---- synthetic inputs ----------------------------------------------------
 * Seeded, platform-independent generators for the benchmark scenes. */
/* 4-octave Perlin fBm density clamped to [0,1] on res^3 (x fastest). */
ne_status ne_gen_fbm_grid(int32_t res, uint64_t seed, float base_freq, float* out);
/* Cumulus-like density: (Worley F1 - fBm) shaped by an ellipsoid, in [0,1]. */
ne_status ne_gen_cumulus_grid(int32_t res, uint64_t seed, float* out);
/* Icosphere subdivided `subdiv` times (20*4^subdiv triangles) with vertices
 * displaced along the normal by amp * fBm(seed); unit radius.  verts gets
 * n_vertices*3 floats, tris n_triangles*3 indices; pass NULL to query sizes. */
ne_status ne_gen_icosphere(int32_t subdiv, uint64_t seed, float amp, float* verts, uint32_t* tris,
                           uint32_t* n_vertices, uint32_t* n_triangles);

/* ---- scene files -----------------------------------------------------------
 * SceneReader::loadScene (src/io/SceneReader.cpp:10-54) for JSON scene files
 * (version 0.0.1a): materials (microfacet, emitter, directionalLight,
 * infiniteAreaLight, volume), primitives (rectangle, sphere, obj, point,
 * volume) with their instance transforms, camera and renderer settings, read
 * into the flattened ne_scene_desc, an ne_camera_desc and ne_scene_settings.
 * Host-only: no context or device is needed.  Errors: NE_ERR_IO for files that
 * cannot be read, NE_ERR_PARSE for malformed JSON, missing members and types
 * the reference rejects with LOG(FATAL), NE_ERR_UNSUPPORTED for valid input
 * this backend cannot represent (image textures, .vdb grids, glTF models,
 * a second volume or environment); the message is in ne_last_error(NULL). */
typedef struct ne_scene_file ne_scene_file;
/* SceneSettings (src/core/Settings.h:9-14) plus the renderer mode. */
typedef struct ne_scene_settings {
  int32_t width, height; /* renderer.resolution */
  int32_t spp;           /* renderer.spp */
  int32_t bounces;       /* renderer.bounces */
  int32_t hdr;           /* renderer.HDR */
  int32_t realtime;      /* renderer.mode == "realtime" */
} ne_scene_settings;
enum {
  /* Substitute what the backend cannot represent instead of failing, and list
   * each substitution in ne_scene_file_warnings: image albedo -> grey 0.5,
   * OBJ without materialName -> glossy grey 0.5 roughness 0.65, non-colliding
   * emitter sphere -> colliding, constant normal maps ignored. */
  NE_LOAD_LENIENT = 1
};
typedef struct ne_load_options {
  int32_t flags;             /* NE_LOAD_* */
  int32_t maj_res;           /* majorant cells per axis; 0: 16, capped at the largest grid dimension */
  const char* resources_dir; /* base of relative asset paths (the reference's RESOURCES_DIR);
                                NULL: the scene file's directory (the working directory for parse) */
} ne_load_options;
/* Reads and flattens the scene file at `path` (opt may be NULL). */
ne_status ne_scene_file_read(const char* path, const ne_load_options* opt, ne_scene_file** out);
/* Same from JSON text in memory. */
ne_status ne_scene_file_parse(const char* json, size_t length, const ne_load_options* opt, ne_scene_file** out);
/* The flattened scene, the main camera (up fixed to (0,1,0) and aspect
 * width/height, as SceneReader.cpp:612) and the settings.  Pointers in *desc
 * stay valid until the file is released.  Any output may be NULL. */
ne_status ne_scene_file_get(const ne_scene_file* file, ne_scene_desc* desc, ne_camera_desc* camera,
                            ne_scene_settings* settings);
/* Substitutions made under NE_LOAD_LENIENT, one per line ("" if none). */
const char* ne_scene_file_warnings(const ne_scene_file* file);
void ne_scene_file_release(ne_scene_file* file);
/* Text grid reader, ResourceManager::loadVolasTexture
 * (src/core/ResourceManager.cpp:222-286): line 1 "W H D", line 2 ignored,
 * then W*H*D whitespace-separated floats, x fastest.  grid == NULL queries the
 * resolution only; otherwise capacity (floats) must hold W*H*D. */
ne_status ne_vol_read(const char* path, int32_t res[3], float* grid, uint64_t capacity);

/* ---- output images ----------------------------------------------------------
 * Writers for width x height RGB float images (row-major, x fastest, row 0 at
 * the top), host-only:
 *   NE_IMAGE_PPM16  OfflineEngine::coreLoop's output.ppm
 *                   (src/core/OfflineEngine.cpp:80-137): P6, maxval 65535,
 *                   pixels from the last index to the first, channel values
 *                   clamped to [0, 1] and truncated from value * 65535
 *   NE_IMAGE_PNG    saveImage for PNG (src/materials/Texture.h:44-60): 8-bit
 *                   RGB, clamp to [0, 1] then value * 255 truncated
 *   NE_IMAGE_EXR    saveImage for EXR (Texture.h:62-69): scanline OpenEXR,
 *                   FP32 channels, no compression
 *   NE_IMAGE_PFM    linear Portable FloatMap (rows stored bottom to top)
 * Post-processed pixels (ne_film_tonemap) suit PPM16 and PNG; linear sums
 * divided by spp suit EXR and PFM. */
enum { NE_IMAGE_PPM16 = 0, NE_IMAGE_PNG = 1, NE_IMAGE_EXR = 2, NE_IMAGE_PFM = 3 };
ne_status ne_image_write(const char* path, int32_t format, int32_t width, int32_t height, const float* rgb);

#ifdef __cplusplus
}
#endif
#endif /* NE_GPU_H */
