// Host-side internals shared by the .hip translation units of libsr_gicp.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstring>
#include <string>
#include <vector>

#include "sr_device.hpp"

namespace sr {

// Error raised by host code; carries the SpatialError status code.
struct Error {
  int32_t code;
  std::string msg;
};

[[noreturn]] void throw_hip(hipError_t e, const char* what);
#define SR_HIP(call)                               \
  do {                                             \
    hipError_t e_ = (call);                        \
    if (e_ != hipSuccess) ::sr::throw_hip(e_, #call); \
  } while (0)

[[noreturn]] inline void fail(int32_t code, const std::string& msg) { throw Error{code, msg}; }

// ---- device memory -------------------------------------------------------

// Caching allocator: freed blocks go to a size-bucketed free list and are reused
// by later calls of the same process (hipMalloc/hipFree are far too slow to sit
// on the per-registration path).
void* dev_alloc(size_t bytes);
void dev_free(void* p);
// True while the calling thread runs the allocation-free entry point: its
// buffers then come from the library's static scratch array.
bool arena_active();
// Hands device memory that kernels queued on stream s may still read to the
// calling thread's next download() on s (which waits for the stream), which
// releases it; lets a host function return without waiting for its kernels.
void release_after_sync(void* p, hipStream_t s);

// RAII buffer from the caching allocator.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  bool owned = true;  // false: a view into memory another buffer owns
  DevBuf() = default;
  explicit DevBuf(size_t count) { resize(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), owned(o.owned) {
    o.p = nullptr;
    o.n = 0;
    o.owned = true;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      owned = o.owned;
      o.p = nullptr;
      o.n = 0;
      o.owned = true;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  // A non-owning view of count elements at ptr (the owner must outlive it).
  static DevBuf view(T* ptr, size_t count) {
    DevBuf b;
    b.p = ptr;
    b.n = count;
    b.owned = false;
    return b;
  }
  void resize(size_t count) {
    if (count <= n && p) return;
    release();
    p = static_cast<T*>(dev_alloc(count * sizeof(T) > 0 ? count * sizeof(T) : 16));
    n = count;
  }
  // Leaves the memory to release_after_sync (see there); false inside the
  // allocation-free entry point, whose caller must then wait for s itself.
  bool release_after(hipStream_t s) {
    if (arena_active()) return false;
    if (p && owned) release_after_sync(p, s);
    p = nullptr;
    n = 0;
    return true;
  }
  void release() {
    if (p && owned) dev_free(p);
    p = nullptr;
    n = 0;
    owned = true;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Per-thread stream (default priority) of the calling host thread.
hipStream_t thread_stream();

// Host arrays packed for one upload: add() returns each array's byte offset
// (256-byte aligned), at() the device pointer at an offset of the uploaded copy.
struct HostPack {
  std::vector<uint8_t> bytes;
  template <typename T>
  size_t add(const std::vector<T>& v) {
    const size_t o = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(o + sizeof(T) * std::max<size_t>(v.size(), 1));
    if (!v.empty()) std::memcpy(bytes.data() + o, v.data(), sizeof(T) * v.size());
    return o;
  }
  template <typename T>
  static const T* at(const uint8_t* base, size_t o) {
    return reinterpret_cast<const T*>(base + o);
  }
};
// hipMemcpyAsync host -> device of a small host buffer (staged in pinned
// memory, so the call does not wait); `src` may be reused once it returns.
void upload(void* dst, const void* src, size_t bytes, hipStream_t s);
// hipMemcpyAsync device -> host of a small buffer, then the stream is
// synchronised (through a pinned buffer, so the copy is a direct DMA).
void download(void* dst, const void* src, size_t bytes, hipStream_t s);
// Ensures the HIP device is initialised; throws SR_STATUS_UNAVAILABLE if not.
void ensure_device();
int device_cu_count();

// ---- kernel timing (sr_profile_*) ------------------------------------------

// When profiling is on, brackets a launch with HIP events on its stream and
// files the pair under `name`. Cheap no-op when off.
struct ProfScope {
  int slot = -1;
  ProfScope(const char* name, hipStream_t s);
  ~ProfScope();
  hipStream_t stream = nullptr;
};

// Device counters for the named search kernel when search statistics are on
// (sr_profile_enable(2)); nullptr otherwise.
SearchStats* prof_stats(const char* name);

// Fills n u32 words with v.
void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);

// ---- scan ----------------------------------------------------------------

// Exclusive prefix sum of n u32 values (in may equal out). If total is non-null
// the grand total is written to *total (device pointer).
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total,
                        hipStream_t s);

// ---- point clouds --------------------------------------------------------

// Device cloud packed as float4 (x, y, z, 0).
struct Cloud {
  DevBuf<float4> pts;
  uint32_t n = 0;
};

// Uploads / packs an xyz view into a float4 device cloud.
void pack_cloud(const float* x, const float* y, const float* z, uint32_t n, bool device_mem,
                Cloud& out, hipStream_t s, uint64_t* h2d_bytes);

// Axis-aligned bounds of a device float4 cloud: writes 6 ordered-u32 words
// (minx,miny,minz,maxx,maxy,maxz) to dev_bounds.
void cloud_bounds(const float4* pts, uint32_t n, uint32_t* dev_bounds, hipStream_t s);

// ---- voxel downsample ----------------------------------------------------

struct VoxelOut {
  Cloud centroids;      // C cells, first-occurrence order
  DevBuf<int64_t> keys; // 3*C (optional)
  DevBuf<uint32_t> counts;
  uint32_t cells = 0;
};

// Source of one segment of pack_many: f32 columns in device memory.
struct SegSource {
  const float* x;
  const float* y;
  const float* z;
  uint32_t n;
  uint32_t offset;  // first point in the packed cloud
};
// Packs many device xyz columns into one float4 cloud, segment after segment.
// With `hold` set, the same sweep also reduces each segment's bounds (6
// ordered words, as seg_bounds) into *dev_bounds, which lives in the buffer
// handed to *hold.
void pack_many(const std::vector<SegSource>& srcs, Cloud& out, hipStream_t s,
               DevBuf<uint8_t>* hold = nullptr, const uint32_t** dev_bounds = nullptr);
// Bounds (6 ordered words per segment) of segments [off[s], off[s+1]) of pts
// into dev_bounds (stream-ordered; the host does not wait).
void seg_bounds(const float4* pts, const std::vector<uint32_t>& off, uint32_t* dev_bounds,
                hipStream_t s);
// voxel_downsample of every segment [off[s], off[s+1]) of `in` (origin = the
// segment's minimum), in one set of launches. The segments' centroids are
// out.centroids[cell_off[s], cell_off[s+1]), each exactly what voxel_downsample
// returns for the segment alone. Synchronises the stream. dev_bounds: the
// segments' bounds when already reduced (pack_many), else computed here.
void voxel_downsample_many(const Cloud& in, const std::vector<uint32_t>& off, float leaf,
                           VoxelOut& out, std::vector<uint32_t>& cell_off, hipStream_t s,
                           const uint32_t* dev_bounds = nullptr);

// Centroid voxel filter of a device cloud (voxel.rs:550-636 semantics).
void voxel_downsample(const Cloud& in, float leaf, const float* origin, bool want_keys,
                      VoxelOut& out, hipStream_t s);

// ---- uniform grid --------------------------------------------------------

struct Grid {
  GridDesc desc{};
  DevBuf<uint32_t> cell_start;  // ncells + 1
  DevBuf<float4> pts;           // grid order; .w = original index (bit cast)
  DevBuf<uint32_t> order;       // grid order -> original index
  DevBuf<uint8_t> empty_radius; // per cell: Chebyshev distance (cells) to the nearest
                                // occupied cell, capped at kEmptyCap + 1 (optional)
  DevBuf<uint32_t> occ;         // occupancy bitmap (desc.bits points here)
  DevBuf<uint32_t> rowocc;      // row bitmap (desc.rows points here)
};

constexpr int kEmptyCap = 14;

// Fills g.empty_radius (Chebyshev distance transform of cell occupancy, capped).
void build_empty_radius(Grid& g, hipStream_t s);
// In place, over the cells of d: field[c] = min over cells c' of
// max(Chebyshev distance(c, c'), field[c']), for a field of 0 (seed cells) and
// kEmptyCap + 1 (the rest): the distance in cells to the nearest seed, capped.
void chebyshev_field(uint8_t* field, const GridDesc& d, hipStream_t s);

// Grids of many segments of one cloud (segment s = points [pt_off[s],
// pt_off[s+1])), each the grid build_grid makes for the segment alone, in
// shared flat arrays: segment s's cell_start is cell_start[cs_off[s] ..] with
// values counting from 0 within it, its points are pts[pt_off[s] ..] (.w = the
// index within the segment), its empty radius is empty_radius[cell_off[s] ..],
// and desc[s].bits / .rows point at its bitmap words.
struct GridSet {
  std::vector<GridDesc> desc;
  std::vector<uint32_t> pt_off, cell_off, cs_off;
  DevBuf<uint32_t> cell_start;
  DevBuf<float4> pts;
  DevBuf<uint32_t> occ, rowocc;
  DevBuf<uint8_t> empty_radius;
};
// choose_cell_edge for every segment (hb: 6 bound words per segment), with the
// occupancy sketches of all segments in one launch; same edges.
void choose_cell_edges_many(const Cloud& c, const std::vector<uint32_t>& off,
                            const std::vector<uint32_t>& hb, float occupancy, float leaf_hint,
                            std::vector<float>& h, hipStream_t s);
// build_grid for every segment, edge h[s], and build_empty_radius for the
// first `empty_segs` segments (the empty radius covers their cells only).
// Synchronises the stream.
void build_grids_many(const Cloud& c, const std::vector<uint32_t>& off,
                      const std::vector<uint32_t>& hb, const std::vector<float>& h,
                      uint32_t empty_segs, GridSet& out, hipStream_t s);

// Chooses a cell edge so that occupied cells hold about `occupancy` points
// (estimated at candidate edges around a guess; leaf_hint > 0 says the cloud
// holds at most one point per voxel of that edge, which centres the guess).
float choose_cell_edge(const Cloud& c, const uint32_t* host_bounds, float occupancy,
                       float leaf_hint = 0.0f);
// Builds a dense counting-sort grid with edge h over a device cloud (points in
// each cell sorted by ascending original index, like build_grid).
void build_grid(const Cloud& c, const uint32_t* host_bounds, float h, Grid& g, hipStream_t s);
// Host copy of cloud bounds (6 ordered words) — synchronises the stream.
void host_bounds(const Cloud& c, uint32_t* host6, hipStream_t s);

// ---- covariances ---------------------------------------------------------

// Largest k of the k-NN kernels (their per-lane sorting networks).
constexpr uint32_t kMaxKnnK = 64;

// Grid-order point records (2 float4 per grid position i, 32 bytes, one
// aligned sector): [2i] = the grid point (x, y, z, original index bits),
// [2i+1] = its plane normal float4(n, a) with C = a*I - (1-eps) n n^T. The
// Gauss-Newton passes gather a matched target point and its normal from one
// record instead of two arrays.
constexpr uint32_t kRec = 2;

// Exact k-NN normals of the grid's own points as grid-order point records
// (a = 1; 2 * npoints float4). If cov6 is non-null also writes the 6
// regularized covariance entries per point in ORIGINAL index order.
void knn_normals(const Grid& g, uint32_t k, double eps, float4* records, float* cov6,
                 float* normal3, hipStream_t s);
// One cloud of a k-NN normals launch over many clouds (outputs as knn_normals).
struct KnnCloud {
  GridDesc g;
  const uint32_t* cs;
  const float4* gp;
  float4* rec;  // grid-order point records (kRec float4 a point)
  float* cov6;
  float* normal3;
  uint32_t block_begin;  // first workgroup of this cloud in the launch (set by the launcher)
  uint32_t pad_;
  // Optional: only the grid positions subset[0 .. *subset_n) are computed
  // (launched for up to g.npoints of them); nrm_orig / tau_orig, when set,
  // receive float4(n, 1) and the k-th neighbour's squared distance in
  // original point order.
  const uint32_t* subset = nullptr;
  const uint32_t* subset_n = nullptr;
  float4* nrm_orig = nullptr;
  float* tau_orig = nullptr;
};
KnnCloud knn_cloud(const Grid& g, float4* records, float* cov6, float* normal3);
// knn_normals over many clouds in one launch (one K for all): a cloud's
// results do not depend on the others.
void knn_normals_many(const std::vector<KnnCloud>& clouds, uint32_t k, double eps, hipStream_t s);
// Radius-grid covariances with COV_GRID_WGSL semantics over a grid of edge =
// radius. normals in grid order of `g_out_order` (the GICP grid); cov6 in original order.
void radius_normals(const Grid& rg, float radius, float eps, float4* normals_orig_order,
                    float* cov6, float* normal3, hipStream_t s);
// Unit normals and curvature (NORMALS_GRID_WGSL semantics) over a grid of
// edge = radius, oriented to `viewpoint` (3 floats, host) when non-null; out in
// original point order as float4(n, curvature).
void grid_normals(const Grid& rg, float radius, const float* viewpoint, float4* out_orig,
                  hipStream_t s);
// Grid-order point records whose normals are per-point float4 values given in
// original order.
void gather_records(const Grid& g, const float4* orig, float4* records, hipStream_t s);

// Exact k-NN of queries against a grid (sr_knn).
void knn_query(const Grid& g, const float4* queries, uint32_t nq, uint32_t k, int32_t* out_idx,
               float* out_d2, hipStream_t s);


// ---- GICP iterations -----------------------------------------------------

// Per-instance device state of the Gauss-Newton loop.
struct InstState {
  double q[4];       // pose rotation (x,y,z,w), FP64
  double t[3];       // pose translation, FP64
  float rf[9];       // f32 rotation matrix used by the per-point pass (high part)
  float tf[3];
  float rlo[9];      // low parts: R = rf + rlo, t = tf + tlo to ~2^-48 relative
  float tlo[3];
  double fitness;
  double last_update;
  uint64_t count;    // correspondences of the last pass
  uint32_t phase;    // 0 iterate, 1 final fitness pass pending, 2 done
  uint32_t iter;     // Gauss-Newton iterations started
  uint32_t status;   // 0 ok, 1 too few correspondences, 2 singular
  uint32_t converged;
  uint32_t counter;  // blocks of the current pass that have finished
  uint32_t pad_;
  // Bound on how far the next pass's pose moves a source point s from where
  // the last pass put it: |dp'| <= move_rot * |s| + move_const (+inf before
  // the first pass, 0 when the pose did not change).
  float move_rot;
  float move_const;
};

struct InstConfig {
  float dmax;       // max_correspondence_distance
  float dmax2;      // dmax*dmax in f32 (gicp.rs:125-126)
  float one_minus_eps;  // 1 - epsilon of the plane regularization
  uint32_t max_iter;
  uint32_t min_corr;
  uint32_t flags;   // kCfgMatchReuse: certified matches may be kept without a search
  double trans_eps;
  double fit_eps;
};

// One registration instance of a pass launch.
struct IterInstance {
  const float4* src_rec;  // source point records (grid order for locality; kRec float4 a point)
  const float4* tgt;      // target points in grid order
  const float4* tgt_rec;  // target point records (grid order, kRec float4 a point)
  const uint32_t* cell_start;
  const uint8_t* empty_radius;  // target grid's empty-radius field (may be null)
  uint32_t* match;        // per source point: target grid position matched by the
                          // previous pass (kNone: none); seeds the next search
  float* budget;          // per source point: how far it may still move before its
                          // match can change (0: search again)
  float4* seen;           // per source point: the transformed point its match was
                          // searched for (x, y, z; NaN before any search)
  float src_radius;       // bound on |s| over the source points
  GridDesc grid;
  InstConfig cfg;
  InstState* state;
  float* partials;        // [nblocks][kTerms][2]
  uint32_t n_src;
  uint32_t block_begin;
  uint32_t nblocks;
  uint32_t ppt;           // source points per thread (from this instance's size only,
                          // so its sums do not depend on the rest of a batch)
};

constexpr uint32_t kCfgMatchReuse = 1u;  // InstConfig.flags

// Library-wide runtime options (sr_option_set / sr_option_get).
struct Options {
  std::atomic<int64_t> match_reuse{1};   // certified match reuse in the Gauss-Newton passes
  std::atomic<int64_t> radius_tiles{0};  // radius covariances / grid normals from LDS cell tiles
  std::atomic<int64_t> autotune{1};      // pick pass launch choices per problem class by timing
  std::atomic<int64_t> tune_force{-1};   // >= 0: use that candidate for every pass loop
  std::atomic<int64_t> odometry_incremental{1};  // odometry map: recompute affected normals only
  std::atomic<int64_t> pass_graphs{1};   // replay repeated pass loops as captured HIP graphs
};
Options& options();

constexpr int kTerms = 28;  // 21 H (upper triangle) + 6 g + fitness sum

// Upper-triangle index of (a, b), a <= b, in a 6x6 symmetric matrix: the order
// of the 21 H terms in a block partial.
__host__ __device__ constexpr int ut(int a, int b) { return a * 6 - a * (a - 1) / 2 + (b - a); }

// A source point whose nearest target point phase 1 of the search could not
// certify, queued for the ring search.
struct HardQuery {
  float px, py, pz;  // transformed source point
  uint32_t inst;     // instance; bit 31 set: phase 1 scanned the 3x3x3 block
  uint32_t i;        // source point within the instance
  uint32_t pos;      // best grid position so far (kNone: none)
  uint64_t key;      // best packed (d2 bits << 32 | original index) so far
};

// Scratch of a pass launch: the queue of uncertified points (room for every
// source point of the launch), its length ctl[0] (zero when a pass starts; the
// finalize kernel zeroes it for the next), one flag per block (a stage wrote
// the block's partial) and four words per block (the first stage's mask of
// the block's points left to search, one bit per thread).
struct PassScratch {
  HardQuery* queue;
  uint32_t* ctl;   // [0]: queue length
  uint32_t* done;
  unsigned long long* pend;
};

// Launch choices of a pass that leave every result unchanged (the tuner picks
// them per problem class): lanes per queued point in the ring search (4, 8 or
// 16), the workgroup cap of that stage (2048 to 8192), and the launch size (blocks) from
// which the first stage runs as two kernels (settle + marked search).
struct PassTuning {
  int group = 8;
  uint32_t hard_cap = 4096;
  uint32_t split_blocks = 4096;
};

// Workgroups of the queue-draining stage for a launch over `points` source
// points, at most `cap`.
uint32_t hard_stage_blocks(uint64_t points, uint32_t cap = 4096);

// Launches one Gauss-Newton pass over `count` instances (device array `insts`),
// total_blocks blocks; block_inst maps block -> instance. Without `finalize`
// only the per-block partial sums are written (no solve, no state change).
void gicp_pass(const IterInstance* insts, const uint32_t* block_inst, uint32_t total_blocks,
               uint32_t count, const PassScratch& ps, uint32_t hard_blocks, hipStream_t s,
               bool finalize = true, const PassTuning& tune = PassTuning{});
// Initialises instance states from poses (device array of 7 doubles per
// instance: quaternion x y z w, translation).
void init_states(InstState* states, const double* init7, uint32_t count, hipStream_t s);
// init_states, plus every source point without a match (kNone), budget (0) or
// searched position (NaN): a batch's starting state in one launch.
void reset_batch_state(InstState* states, const double* init7_dev, uint32_t count, uint32_t* match,
                       float* budget, float4* seen, uint64_t n, hipStream_t s);

// ---- rolling voxel map (odometry) ----------------------------------------

// Rigid transform for map insertion: R row-major and t, FP64.
struct MapPose {
  double r[9];
  double t[3];
};

// Voxel representatives in insertion order plus a key -> list index hash.
struct VoxelMap {
  float leaf = 0.0f, inv_leaf = 0.0f;
  uint64_t capacity = 0;  // list capacity (voxels)
  uint64_t mask = 0;      // hash slots - 1
  DevBuf<uint64_t> keys;
  DevBuf<uint32_t> val, claim, flags, scal;  // scal: [new count, overflow flag]
  Cloud list;             // map points, list.n of them
  uint64_t drops = 0;     // times old voxels were dropped (list indices renumbered)

  void init(float leaf, uint32_t capacity, hipStream_t s);
  // Transforms n scan points (sensor frame) by `pose` and adds the voxels the
  // map lacks; returns the number added. Drops the oldest voxels first when the
  // result would exceed the capacity.
  uint32_t insert(const float4* scan_local, uint32_t n, const MapPose& pose, hipStream_t s);
  void drop_oldest(uint32_t count, hipStream_t s);
};

// Grid positions of g (the grid of the map list's points [0, npoints)) whose
// k-NN normal can differ from the one computed before the points [n_old,
// npoints) were appended: the new points, and every point whose k-th
// neighbour's squared distance ltau (list order, from that computation) is not
// below the squared lower bound on its distance to the new points that the
// cell distance to the nearest new point's cell gives. Every other point keeps
// its k nearest, and visits them in the same order when g has the same
// geometry as before. Writes subset[0 .. *count) (device).
void map_refresh_set(const Grid& g, const float4* list, uint32_t n_old, const float* ltau,
                     uint32_t* subset, uint32_t* count, hipStream_t s);
// Grid-order point records with normal lnrm[g.order[i]] at grid position i.
void gather_list_normals(const Grid& g, const float4* lnrm, float4* records, hipStream_t s);

}  // namespace sr
