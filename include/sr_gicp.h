/*
 * sr_gicp.h — C ABI of the MI355X (gfx950) GICP scan-registration path.
 *
 * This library is the GPU backend behind SpatialRust's GICP call. Each entry
 * point below replaces one reference interface (paths relative to the
 * rsasaki0109/SpatialRust checkout); the replaced interface is cited beside it.
 *
 * Conventions
 *  - Point clouds cross the boundary as borrowed SoA f32 columns (sr_xyz_view),
 *    the same layout as `HasPositions3::positions3`
 *    (crates/spatialrust-core/src/capabilities.rs:22-28). `mem` says whether the
 *    pointers are host (0) or device (1) memory.
 *  - Every entry point returns an int32 status: 0 = Ok, otherwise 1 + the index
 *    of the `SpatialError` variant it maps to
 *    (crates/spatialrust-core/src/error.rs:8-37): 1 MissingField,
 *    2 SchemaValidation, 3 InvalidArgument, 4 Io, 5 BufferLengthMismatch,
 *    6 UnsupportedDType. The message (reference wording where the reference has
 *    one) is available from sr_last_error() on the calling thread.
 *  - SR_STATUS_UNAVAILABLE (100) means "no usable HIP device"; a caller that
 *    asked for ExecutionPolicy::Auto may then run its own CPU path. An explicit
 *    GPU policy gets InvalidArgument instead (crates/spatialrust-core/src/execution.rs:40-48).
 *  - All entry points are thread-safe. One HIP context per process is created
 *    lazily on first use (as `WgpuRuntime::shared`, crates/spatialrust-gpu/src/runtime.rs:90-95).
 *    Calls that share a handle (sr_map, sr_server) must not overlap.
 */
#ifndef SR_GICP_H
#define SR_GICP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SR_OK 0
#define SR_ERR_MISSING_FIELD 1
#define SR_ERR_SCHEMA 2
#define SR_ERR_INVALID_ARGUMENT 3
#define SR_ERR_IO 4
#define SR_ERR_LENGTH_MISMATCH 5
#define SR_ERR_DTYPE 6
#define SR_STATUS_UNAVAILABLE 100

#define SR_MEM_HOST 0u
#define SR_MEM_DEVICE 1u

/* Borrowed xyz columns, valid for the duration of the call. */
typedef struct {
  const float* x;
  const float* y;
  const float* z;
  uint64_t n;
  uint32_t mem; /* SR_MEM_HOST or SR_MEM_DEVICE */
} sr_xyz_view;

/* GicpConfig (crates/spatialrust-registration/src/gicp.rs:23-44) plus the
 * voxel leaf applied to both clouds before registration (<= 0: none).
 * Defaults (gicp.rs:46-60) come from sr_gicp_config_default(). */
typedef struct {
  uint32_t max_iterations;
  float max_correspondence_distance;
  uint32_t k_neighbors;
  double epsilon;
  double transformation_epsilon;
  double fitness_epsilon;
  uint32_t min_correspondences;
  float init_q[4]; /* initial_guess rotation, x y z w */
  float init_t[3]; /* initial_guess translation */
  float covariance_radius; /* <= 0: None (k-NN covariances) */
  float voxel_leaf;        /* <= 0: clouds are used as given */
} sr_gicp_config;

/* ExecutionPolicy / DeviceKind (crates/spatialrust-core/src/execution.rs:5-15,
 * device.rs:4-12). SR_DEV_HIP is the device kind this backend adds. */
typedef enum { SR_CPU_SINGLE = 0, SR_CPU_PARALLEL = 1, SR_GPU = 2, SR_AUTO = 3 } sr_policy_kind;
typedef enum { SR_DEV_CPU = 0, SR_DEV_WGPU = 1, SR_DEV_CUDA = 2, SR_DEV_HIP = 3 } sr_device_kind;

typedef struct {
  int32_t kind;   /* sr_policy_kind */
  int32_t device; /* sr_device_kind, used when kind == SR_GPU */
} sr_policy;

/* RegistrationResult (crates/spatialrust-registration/src/registration.rs:15-24).
 * q/t is the f32 Isometry3; q64/t64 is the FP64 pose the device iterated. */
typedef struct {
  float q[4];
  float t[3];
  double fitness;
  uint64_t iterations;
  uint8_t converged;
  uint8_t pad_[7];
  double q64[4];
  double t64[3];
  uint64_t source_points; /* after the voxel filter */
  uint64_t target_points; /* after the voxel filter */
  uint64_t correspondences; /* in the last Gauss-Newton pass */
} sr_result;

/* Stage bits recorded in sr_receipt.stage_mask (ExecutionReceipt::record_stage,
 * crates/spatialrust-core/src/execution.rs:135-138). */
#define SR_STAGE_VOXEL 0x1u       /* "gicp-voxel" */
#define SR_STAGE_COVARIANCE 0x2u  /* "gicp-covariance" */
#define SR_STAGE_ITERATE 0x4u     /* "gicp-iterate" */
#define SR_STAGE_READBACK 0x8u    /* "gpu-readback" */
#define SR_STAGE_GRID 0x10u       /* "gicp-grid" */

/* ExecutionReceipt (crates/spatialrust-core/src/execution.rs:69-138). */
typedef struct {
  sr_policy requested;
  sr_policy resolved;
  uint64_t h2d_bytes;
  uint64_t d2d_bytes;
  uint64_t d2h_bytes;
  uint32_t stage_mask;
  uint32_t pad_;
} sr_receipt;

/* ---- runtime ---------------------------------------------------------- */

/* Thread-local message of the last failed call on this thread ("" if none). */
const char* sr_last_error(void);

/* 1 if a gfx950 HIP device is usable, else 0 (WgpuRuntime::shared().is_ok(),
 * crates/spatialrust-gpu/src/runtime.rs:90-95). */
int32_t sr_device_available(void);

/* Library ABI version (major*10000 + minor*100 + patch). */
int32_t sr_version(void);

/* GicpConfig::default() (crates/spatialrust-registration/src/gicp.rs:46-60). */
void sr_gicp_config_default(sr_gicp_config* out);

/* ---- runtime options -----------------------------------------------------
 * No counterpart in the reference (its GICP has no device-side state to
 * tune). Process-wide switches read when a registration is set up:
 *   "match_reuse" (default 1): a Gauss-Newton pass may keep a source point's
 *     correspondence without searching when the last search proved that no
 *     other target point can have become nearer (half the gap to the
 *     second-nearest point exceeds the pose's movement). 0 searches every
 *     point in every pass; results are the same either way.
 *   "radius_tiles" (default 0): 1 makes the radius covariances and grid
 *     normals read each point's 27-cell neighbourhood from LDS-staged cell
 *     tiles instead of from device memory per point (slower on gfx950, where
 *     the per-point reads hit L2). Results are bit-identical either way.
 *   "autotune" (default 1): the launch choices of the Gauss-Newton pass loop
 *     (lanes per queued point in the ring search, that stage's workgroup cap,
 *     the launch size from which the first stage splits in two kernels) are
 *     picked per problem class (architecture; one or many registrations; log2
 *     of the source points) by timing each candidate on the class's first
 *     loop, which then costs a few loops (one registration under 128k source
 *     points is not tuned); 0 uses the defaults. Results are
 *     bit-identical whatever is picked. SR_TUNE_CACHE (environment) names a
 *     file the decisions are read from and appended to.
 *   "tune_force" (default -1): >= 0 uses that candidate for every loop (read
 *     "tune_candidates_single" / "tune_candidates_batch" for their counts).
 *   "odometry_incremental" (default 1): after a scan is merged, the odometry
 *     map recomputes the k-NN normals only of the new voxels and of the voxels
 *     whose k-th nearest neighbour is not provably closer than every new voxel
 *     (from their grid cells' distance to the new voxels' cells); 0 recomputes
 *     every normal. Poses are bit-identical either way.
 *   "pass_graphs" (default 1): a pass loop whose launch shape repeats is
 *     captured as a HIP graph on its second run and replayed after that; 0
 *     keeps every loop on eager launches (for external profilers). Results
 *     are bit-identical either way.
 * Unknown names and values out of range are InvalidArgument. */
int32_t sr_option_set(const char* name, int64_t value);
int32_t sr_option_get(const char* name, int64_t* value);
/* The tuner's decisions so far as text, one line per problem class (the
 * chosen launch choices and each candidate's measured loop time in ms); at
 * most cap - 1 bytes plus a NUL go to out, *len receives the full length. */
int32_t sr_tune_describe(char* out, uint64_t cap, uint64_t* len);

/* ---- kernel timing ------------------------------------------------------
 * No counterpart in the reference (it records no GPU timestamps,
 * `timestamp_writes: None`). When enabled, launches of the named kernels are
 * bracketed with HIP events on the stream they run on. Names: "gicp_pass",
 * "knn_normals", "radius_cov", "voxel_insert". */
void sr_profile_enable(int32_t on);
void sr_profile_reset(void);
/* Sum of event-measured durations (ms) and number of launches of `name`. */
int32_t sr_profile_read(const char* name, double* total_ms, uint64_t* launches);
/* Individual durations (ms) of up to `cap` launches of `name`, in launch order;
 * returns the number recorded (or -1 on a HIP error). */
int64_t sr_profile_read_each(const char* name, double* out_ms, uint64_t cap);
/* With sr_profile_enable(2), the grid searches also count their work; reads and
 * clears {queries, rings, cells (rows) visited, points tested, then queries
 * rejected without a scan, finished in the 3x3x3 block, finished in the ring
 * search with a neighbour, finished there without one, then the largest cell
 * and point counts of a single query} of "knn_normals" or "gicp_pass". */
int32_t sr_profile_read_stats(const char* name, uint64_t* out10);

/* ---- GICP (drop-in for GicpRegistration::align_with_diagnostics) ------- */

/* Replaces GicpRegistration::align / align_with_diagnostics
 * (crates/spatialrust-registration/src/gicp.rs:103-224, 232-234) and adds the
 * policy/receipt surface its sibling operators already have
 * (crates/spatialrust-filtering/src/voxel.rs:213-237). `receipt` may be NULL. */
int32_t sr_gicp_align(const sr_xyz_view* source, const sr_xyz_view* target,
                      const sr_gicp_config* config, sr_policy policy,
                      sr_result* out, sr_receipt* receipt);

/* sr_gicp_align without device allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * semantics and bit-identical results as sr_gicp_align with policy
 * Gpu(Hip); every device buffer of the call comes from a 256 MiB scratch array
 * the library holds on each GPU it runs on until the process ends (query it
 * with the read-only options "noalloc_scratch_bytes" and "noalloc_peak_bytes").
 * Each cloud may hold at most 2,000,000 points; a call whose buffers would
 * not fit the scratch fails with InvalidArgument. Calls on one GPU share its
 * array and are serialised (a second caller waits); no graph is captured. */
int32_t sr_gicp_align_noalloc(const sr_xyz_view* source, const sr_xyz_view* target,
                              const sr_gicp_config* config, sr_result* out);

/* One correspondence-and-accumulate pass of the Gauss-Newton loop at a given
 * FP64 pose (q = x y z w, t), without the solve: the loop body of
 * GicpRegistration::align_with_diagnostics (crates/spatialrust-registration/src/gicp.rs:137-179)
 * as the device runs it. Per source point, the index of its target
 * correspondence (the nearest target point within max_correspondence_distance,
 * ties to the lower index) or -1; the FP64 totals H (6x6 row-major, full
 * symmetric), g (6), the correspondence count and the sum of squared
 * correspondence distances. Clouds are used as given (voxel_leaf is ignored);
 * covariances follow config (k_neighbors or covariance_radius). Any output may
 * be NULL. For checking the device pass against a CPU evaluation. */
int32_t sr_gicp_normal_equations(const sr_xyz_view* source, const sr_xyz_view* target,
                                 const sr_gicp_config* config, const double* q, const double* t,
                                 int32_t* out_match, double* out_H, double* out_g,
                                 uint64_t* out_count, double* out_d2_sum);

/* ---- voxel downsample (VoxelGridDownsample, centroid mode, xyz schema) -- */

/* Replaces VoxelGridDownsample::filter_with_policy for the centroid/xyz case
 * (crates/spatialrust-filtering/src/voxel.rs:213-237, 251-368, 550-636).
 * `origin` NULL = per-axis minimum (VoxelGridDownsampleConfig::centroid,
 * voxel.rs:138-146). Outputs are host arrays of capacity >= input n; cells come
 * in first-occurrence order. `keys` (3 x int64 per cell) and `counts` may be
 * NULL. *n_out receives the cell count. */
int32_t sr_voxel_downsample(const sr_xyz_view* input, float leaf, const float* origin,
                            sr_policy policy, float* out_x, float* out_y, float* out_z,
                            int64_t* keys, uint32_t* counts, uint64_t* n_out,
                            sr_receipt* receipt);

/* ---- per-point plane covariances / normals ----------------------------- */

/* Replaces `covariances` + `regularize` (crates/spatialrust-registration/src/gicp.rs:259-291)
 * when radius <= 0 (exact k-NN, k including the point itself), and
 * estimate_plane_covariances_grid_gpu (crates/spatialrust-gpu/src/kernels/covariances_grid.rs:216-364)
 * when radius > 0. Output: 6 f32 per point [c00,c11,c22,c01,c02,c12]
 * (GpuCovariance, covariances_grid.rs:21-23) and optionally the unit normal
 * (smallest-eigenvalue eigenvector, 3 f32 per point; NULL to skip). Host arrays. */
int32_t sr_plane_covariances(const sr_xyz_view* cloud, uint32_t k_neighbors, float radius,
                             double epsilon, float* out_cov6, float* out_normals);

/* Replaces GpuNormalEstimator::estimate with NormalEstimationConfig::search_radius
 * = Some(radius) (crates/spatialrust-features/src/normal_gpu.rs:33-82): per point
 * the unit normal of the smallest eigenvalue of the radius-neighbourhood
 * covariance and its curvature lambda_min / trace, with NORMALS_GRID_WGSL
 * semantics (crates/spatialrust-gpu/src/kernels/normals_grid.rs:25-198: grid of
 * edge = radius, fewer than 3 neighbours give (0,0,1) and curvature 0), then
 * oriented toward `viewpoint` (3 floats, may be NULL) as
 * orient_normal_towards_viewpoint does (crates/spatialrust-features/src/normal.rs:406-421).
 * Host outputs: out_normals 3 f32 per point, out_curvature 1 f32 per point
 * (either may be NULL). */
int32_t sr_grid_normals(const sr_xyz_view* cloud, float radius, const float* viewpoint,
                        float* out_normals, float* out_curvature);

/* ---- nearest neighbour (NearestNeighborIndex::nearest_k / nearest_one) - */

/* Exact k-NN of each query against `points` on a device uniform grid.
 * Replaces KdTree::nearest_k (crates/spatialrust-search/src/kdtree.rs:306-346)
 * for batches of queries. Neighbours of each query are sorted by ascending
 * squared distance, ties by lower index (brute_force_knn order,
 * crates/spatialrust-search/src/brute.rs:84-110). Missing slots (fewer than
 * k points) get index -1. Host arrays: out_index int32[nq*k], out_d2 f32[nq*k]. */
int32_t sr_knn(const sr_xyz_view* points, const sr_xyz_view* queries, uint32_t k,
               int32_t* out_index, float* out_d2);

/* ---- device-resident target maps --------------------------------------- */

typedef struct sr_map sr_map;

/* Uploads and preprocesses a target cloud once (voxel filter per
 * config->voxel_leaf, uniform grid, plane covariances) so many registrations can
 * reuse it. This is the target half of align_with_diagnostics
 * (gicp.rs:119-123) made resident. */
int32_t sr_map_create(const sr_xyz_view* target, const sr_gicp_config* config, sr_map** out);
void sr_map_destroy(sr_map* map);
/* Number of target points kept after the voxel filter. */
uint64_t sr_map_points(const sr_map* map);
/* Device bytes held by the map. */
uint64_t sr_map_device_bytes(const sr_map* map);

/* Registers one source cloud against a resident map (same semantics as
 * sr_gicp_align with the map's target). */
int32_t sr_map_align(const sr_map* map, const sr_xyz_view* source, const sr_gicp_config* config,
                     sr_result* out);

/* ---- batched registration ---------------------------------------------- */

/* Registers `count` independent (source, map) pairs in one set of launches
 * (one launch per Gauss-Newton pass for the whole batch).
 * Each result equals what sr_map_align returns for that pair alone. The
 * reference registers one pair per call; this is the batched form of
 * GicpRegistration::align (gicp.rs:103-224). */
int32_t sr_batch_align(const sr_map* const* maps, const sr_xyz_view* sources, uint32_t count,
                       const sr_gicp_config* config, sr_result* out);

/* Registers `count` independent (source, target) pairs: every cloud is
 * preprocessed (voxel filter, grid, covariances) on a pool of worker streams,
 * then one set of launches runs the Gauss-Newton passes of all pairs. Each
 * result equals sr_gicp_align of that pair alone. The batched form of
 * GicpRegistration::align (crates/spatialrust-registration/src/gicp.rs:103-224). */
int32_t sr_batch_align_pairs(const sr_xyz_view* sources, const sr_xyz_view* targets,
                             uint32_t count, const sr_gicp_config* config, sr_result* out);

/* ---- continuous-batching server ---------------------------------------- */

/* A serving loop over resident maps. The reference has no counterpart (it
 * registers one pair per call, gicp.rs:103-224); every request it serves is
 * that call. Requests queue on the host; each sr_server_step admits queued
 * requests (FIFO) while active slots and the HBM budget allow, runs one
 * Gauss-Newton pass over every active request in one launch, preprocesses the
 * newly admitted sources on worker streams meanwhile, and retires finished
 * requests. A request's result is bit-identical to sr_map_align of its source
 * against its map with the request config (gicp + max_iterations and
 * delta_pose_epsilon as transformation_epsilon). Submit and poll may be called
 * from other threads while a step runs; steps must not overlap. A step does
 * not wait for its own pass: the states return asynchronously and a later
 * step retires what they show finished, with at most two steps in flight. */
typedef struct sr_server sr_server;

typedef struct {
  uint64_t hbm_budget_bytes; /* admission limit for per-request device memory */
  uint32_t max_active;       /* slots: requests iterating together */
  uint32_t max_iterations;   /* per request (0: the gicp config's) */
  double delta_pose_epsilon; /* converged when ||dx|| < this (<= 0: the gicp config's) */
} sr_server_config;

int32_t sr_server_create(const sr_server_config* config, const sr_gicp_config* gicp,
                         sr_server** out);
void sr_server_destroy(sr_server* server);
/* Queues a registration of `source` (host xyz, copied) against `map`; returns a
 * ticket in *ticket. A request whose device footprint alone exceeds the budget
 * is refused (InvalidArgument). */
int32_t sr_server_submit(sr_server* server, const sr_map* map, const sr_xyz_view* source,
                         uint64_t* ticket);
/* One scheduling step (see above). *pending receives the number of requests
 * still queued or active. */
int32_t sr_server_step(sr_server* server, uint32_t* pending);
/* *ready = 1 and *out filled when `ticket` has finished (its error, if any, is
 * returned as the status); the ticket is then released. */
int32_t sr_server_poll(sr_server* server, uint64_t ticket, sr_result* out, uint8_t* ready);
/* Seconds from submit to the step that finished `ticket` (before polling it). */
int32_t sr_server_latency(sr_server* server, uint64_t ticket, double* seconds);
/* Lists up to `cap` finished tickets not listed before (in finishing order;
 * *n receives the count), so a caller polls only finished requests. */
int32_t sr_server_finished(sr_server* server, uint64_t* tickets, uint64_t cap, uint64_t* n);
/* Device bytes currently charged against the budget: for a queued or
 * preparing request an estimate of its preparation peak, for an active one
 * the measured footprint of its prepared source, state and pass buffers. */
uint64_t sr_server_device_bytes(const sr_server* server);

/* ---- scan-to-map odometry (rolling device-resident voxel map) ------------ */

/* A GICP scan matcher over a rolling map. The reference's odometry contract is
 * the ScanMatcher trait (crates/spatialrust-mapping/src/scan.rs:12-20), which it
 * implements only with ICP and scan-to-scan; this keeps the map on the device:
 * each pushed scan is voxel-filtered (voxel.rs:550-636 semantics, origin = the
 * scan's minimum), registered against the map from a constant-velocity guess
 * (the last pose composed with the last motion), and its voxels absent from
 * the map (keys floor(p / leaf) in the map frame) are appended; the lowest
 * point index of a scan wins a new voxel. Beyond max_map_points the oldest
 * voxels are dropped. The first scan defines the map frame. */
typedef struct sr_odometry sr_odometry;

typedef struct {
  float voxel_leaf;        /* scan filter and map voxel edge */
  uint32_t pad_;
  uint64_t max_map_points; /* rolling map capacity (voxels) */
} sr_odometry_config;

int32_t sr_odometry_create(const sr_odometry_config* config, const sr_gicp_config* gicp,
                           sr_odometry** out);
void sr_odometry_destroy(sr_odometry* odometry);
/* Registers and merges one scan (sensor-frame xyz). out->q64/t64 (and q/t) is
 * map_T_scan; source_points = scan voxels, target_points = map voxels after
 * the merge. */
int32_t sr_odometry_push(sr_odometry* odometry, const sr_xyz_view* scan, sr_result* out);
/* sr_odometry_push for one scan of each of `count` distinct odometries (one
 * stream each, e.g. one vehicle each): the scans are filtered and prepared
 * concurrently, registered in one batch (a set of pass launches over all of
 * them) and merged into their maps concurrently. out[i] (optional) is what
 * sr_odometry_push(odometries[i], &scans[i], ...) would return, bit for bit.
 * A registration error fails the call, naming the scan, before any map
 * changes. No counterpart in the reference (ScanMatcher matches one scan per
 * call, crates/spatialrust-mapping/src/scan.rs:12-20). */
int32_t sr_odometry_push_many(sr_odometry* const* odometries, const sr_xyz_view* scans,
                              uint32_t count, sr_result* out);
uint64_t sr_odometry_map_points(const sr_odometry* odometry);
/* Map upkeep counters (no counterpart in the reference): out4 = {map points,
 * normals computed by the last push, full grid rebuilds so far, scans}. The
 * map's grid keeps its geometry from one full rebuild to the next; a rebuild
 * follows a drop of old voxels, a new voxel outside the grid's box, or the map
 * doubling since the last rebuild. */
int32_t sr_odometry_map_stats(const sr_odometry* odometry, uint64_t* out4);
/* Copies up to `cap` map points (interleaved xyz, insertion order) to host
 * memory; *n receives the map size. */
int32_t sr_odometry_map_read(const sr_odometry* odometry, float* xyz, uint64_t cap, uint64_t* n);

#ifdef __cplusplus
}
#endif

#endif /* SR_GICP_H */
