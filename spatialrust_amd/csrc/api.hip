// C ABI of libsr_gicp.so (include/sr_gicp.h) and the host runtime behind it:
// lazy device context, per-thread streams, a caching device allocator, target
// maps, single / batched registration and the continuous-batching server.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <exception>
#include <fstream>
#include <functional>
#include <thread>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/sr_gicp.h"
#include "internal.hpp"

namespace sr {

// ------------------------------------------------------------- runtime ----

static thread_local std::string t_last_error;

[[noreturn]] void throw_hip(hipError_t e, const char* what) {
  throw Error{SR_ERR_INVALID_ARGUMENT, std::string("HIP error in ") + what + ": " + hipGetErrorString(e)};
}

struct DeviceState {
  int status = -1;  // 0 usable, 1 unavailable
  std::string reason;
  int cus = 0;
};

static DeviceState& device_state() {
  static DeviceState ds;
  static std::once_flag once;
  std::call_once(once, [] {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
      ds.status = 1;
      ds.reason = "no HIP device is visible";
      return;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
      ds.status = 1;
      ds.reason = "hipGetDeviceProperties failed";
      return;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      ds.status = 1;
      ds.reason = std::string("device 0 is ") + prop.gcnArchName + ", this library is built for gfx950";
      return;
    }
    if (hipSetDevice(0) != hipSuccess) {
      ds.status = 1;
      ds.reason = "hipSetDevice(0) failed";
      return;
    }
    ds.cus = prop.multiProcessorCount;
    ds.status = 0;
  });
  return ds;
}

void ensure_device() {
  DeviceState& ds = device_state();
  if (ds.status != 0) throw Error{SR_STATUS_UNAVAILABLE, "HIP device unavailable: " + ds.reason};
  static thread_local bool bound = false;
  if (!bound) {
    SR_HIP(hipSetDevice(0));
    bound = true;
  }
}

int device_cu_count() { return device_state().cus; }

hipStream_t thread_stream() {
  struct Holder {
    hipStream_t s = nullptr;
    ~Holder() {
      if (s) (void)hipStreamDestroy(s);
    }
  };
  static thread_local Holder h;
  if (!h.s) SR_HIP(hipStreamCreateWithFlags(&h.s, hipStreamNonBlocking));
  return h.s;
}

// Caching allocator: sizes rounded up to a power of two (>= 512 B); freed
// blocks are kept per size class. Blocks are handed back only to the thread
// that freed them, whose stream orders every later use after earlier work.
struct Pool {
  std::mutex mu;
  std::unordered_map<void*, size_t> live;
  std::map<std::pair<size_t, uint64_t>, std::vector<void*>> free_by_owner;
  uint64_t bytes_live = 0;
};
static Pool& pool() {
  static Pool* p = new Pool();  // never destroyed: blocks stay valid until exit
  return *p;
}
static uint64_t thread_token() {
  static std::atomic<uint64_t> next{1};
  static thread_local uint64_t tok = next++;
  return tok;
}

static size_t size_class(size_t bytes) {
  size_t c = 512;
  while (c < bytes) c <<= 1;
  return c;
}

// Arena of the allocation-free entry point (sr_gicp_align_noalloc): while a
// thread has one installed, its device buffers come from that arena instead
// of the caching allocator.
struct Arena {
  char* base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_;           // offset -> length, coalesced
  std::unordered_map<void*, size_t> live;   // pointer -> length
  size_t in_use = 0, peak = 0;

  void reset(char* b, size_t n) {
    base = b;
    size = n;
    free_.clear();
    free_[0] = n;
    live.clear();
    in_use = peak = 0;
  }
  // First fit, 256-byte granules.
  void* alloc(size_t bytes) {
    const size_t len = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < len) continue;
      const size_t off = it->first, have = it->second;
      free_.erase(it);
      if (have > len) free_[off + len] = have - len;
      void* p = base + off;
      live[p] = len;
      in_use += len;
      peak = std::max(peak, in_use);
      return p;
    }
    fail(SR_ERR_INVALID_ARGUMENT,
         "the buffers of this call need more than the " + std::to_string(size >> 20) +
             " MiB scratch of the allocation-free entry point (" + std::to_string(in_use) +
             " bytes held, " + std::to_string(len) + " more requested)");
  }
  // false: p is not an arena block.
  bool release(void* p) {
    auto it = live.find(p);
    if (it == live.end()) return false;
    size_t off = size_t(static_cast<char*>(p) - base), len = it->second;
    live.erase(it);
    in_use -= len;
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && off + len == nx->first) {
      len += nx->second;
      nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {
      auto pv = std::prev(nx);
      if (pv->first + pv->second == off) {
        pv->second += len;
        return true;
      }
    }
    free_[off] = len;
    return true;
  }
};
static thread_local Arena* t_arena = nullptr;

bool arena_active() { return t_arena != nullptr; }

void* dev_alloc(size_t bytes) {
  if (t_arena) return t_arena->alloc(bytes);
  Pool& P = pool();
  const size_t cls = size_class(bytes);
  const uint64_t tok = thread_token();
  {
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.free_by_owner.find({cls, tok});
    if (it != P.free_by_owner.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      P.live[p] = cls;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, cls);
  if (e != hipSuccess) {
    // Give cached blocks back to the driver and retry once.
    SR_HIP(hipDeviceSynchronize());
    {
      std::lock_guard<std::mutex> lk(P.mu);
      for (auto& kv : P.free_by_owner)
        for (void* q : kv.second) (void)hipFree(q);
      P.free_by_owner.clear();
    }
    (void)hipGetLastError();
    e = hipMalloc(&p, cls);
    if (e != hipSuccess) throw_hip(e, "hipMalloc");
  }
  std::lock_guard<std::mutex> lk(P.mu);
  P.live[p] = cls;
  P.bytes_live += cls;
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  if (t_arena && t_arena->release(p)) return;
  Pool& P = pool();
  std::lock_guard<std::mutex> lk(P.mu);
  auto it = P.live.find(p);
  if (it == P.live.end()) return;
  size_t cls = it->second;
  P.live.erase(it);
  P.free_by_owner[{cls, thread_token()}].push_back(p);
}

// Small host-to-device copies (descriptor tables, offsets) go through a
// per-thread ring of pinned staging chunks: from pinned memory hipMemcpyAsync
// is asynchronous, while from pageable memory it stages the data and the host
// waits ~10 us per call. A chunk is written again only after an event recorded
// behind its last copy has completed. Inside the allocation-free entry point
// (which must not allocate pinned memory either) and for large copies the
// source is used directly.
namespace {
struct Staging {
  static constexpr size_t kChunk = size_t(1) << 20;
  static constexpr int kChunks = 4;
  struct Chunk {
    char* p = nullptr;
    hipEvent_t ev = nullptr;
    hipStream_t s = nullptr;  // stream of the copies from the chunk's current round
    bool pending = false;     // ev recorded behind them and not yet waited for
  };
  Chunk c[kChunks];
  int cur = 0;
  size_t used = 0;
  ~Staging() {
    for (Chunk& ch : c) {
      if (ch.ev) (void)hipEventDestroy(ch.ev);
      if (ch.p) (void)hipHostFree(ch.p);
    }
  }
  Chunk& open(int i) {
    Chunk& ch = c[i];
    if (!ch.p) {
      SR_HIP(hipHostMalloc(reinterpret_cast<void**>(&ch.p), kChunk, hipHostMallocDefault));
      SR_HIP(hipEventCreateWithFlags(&ch.ev, hipEventDisableTiming));
    } else if (ch.pending) {
      SR_HIP(hipEventSynchronize(ch.ev));
      ch.pending = false;
    }
    ch.s = nullptr;
    return ch;
  }
};
}  // namespace

void upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  if (t_arena || bytes > Staging::kChunk / 4) {
    SR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return;
  }
  static thread_local Staging st;
  Staging::Chunk* ch = &st.c[st.cur];
  if (!ch->p) ch = &st.open(st.cur);
  const size_t len = (bytes + 255) & ~size_t(255);
  if (st.used + len > Staging::kChunk || (ch->s && ch->s != s)) {
    if (ch->s) {  // close this chunk's round behind its last copy
      SR_HIP(hipEventRecord(ch->ev, ch->s));
      ch->pending = true;
    }
    st.cur = (st.cur + 1) % Staging::kChunks;
    ch = &st.open(st.cur);
    st.used = 0;
  }
  char* p = ch->p + st.used;
  st.used += len;
  std::memcpy(p, src, bytes);
  ch->s = s;
  SR_HIP(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, s));
}

// A small device-to-host read that the host waits for: through a per-thread
// pinned buffer (a direct DMA), then copied out; the stream is synchronised.
// Buffers waiting for a download() on their stream (release_after_sync).
static thread_local std::vector<std::pair<void*, hipStream_t>> t_after_sync;

void release_after_sync(void* p, hipStream_t s) { t_after_sync.emplace_back(p, s); }

// After a wait for stream s: release what was left to it.
static void released_by_sync(hipStream_t s) {
  auto& v = t_after_sync;
  size_t k = 0;
  for (auto& e : v) {
    if (e.second == s) dev_free(e.first);
    else v[k++] = e;
  }
  v.resize(k);
}

void download(void* dst, const void* src, size_t bytes, hipStream_t s) {
  constexpr size_t kPinned = size_t(256) << 10;
  if (bytes == 0) {
    SR_HIP(hipStreamSynchronize(s));
    released_by_sync(s);
    return;
  }
  if (t_arena || bytes > kPinned) {
    SR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    SR_HIP(hipStreamSynchronize(s));
    released_by_sync(s);
    return;
  }
  struct Pinned {
    void* p = nullptr;
    ~Pinned() {
      if (p) (void)hipHostFree(p);
    }
  };
  static thread_local Pinned pb;
  if (!pb.p) SR_HIP(hipHostMalloc(&pb.p, kPinned, hipHostMallocDefault));
  SR_HIP(hipMemcpyAsync(pb.p, src, bytes, hipMemcpyDeviceToHost, s));
  SR_HIP(hipStreamSynchronize(s));
  released_by_sync(s);
  std::memcpy(dst, pb.p, bytes);
}

// --------------------------------------------------------------- timing ----

struct Profiler {
  std::mutex mu;
  std::atomic<int> on{0};
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> spare;
  std::map<std::string, SearchStats*> stats;  // device counters (search-statistics mode)
};
static Profiler& profiler() {
  static Profiler* p = new Profiler();
  return *p;
}

static hipEvent_t prof_event() {
  Profiler& P = profiler();
  {
    std::lock_guard<std::mutex> lk(P.mu);
    if (!P.spare.empty()) {
      hipEvent_t e = P.spare.back();
      P.spare.pop_back();
      return e;
    }
  }
  hipEvent_t e;
  SR_HIP(hipEventCreate(&e));
  return e;
}

SearchStats* prof_stats(const char* name) {
  Profiler& P = profiler();
  if (P.on.load() != 2 || t_arena) return nullptr;
  std::lock_guard<std::mutex> lk(P.mu);
  auto it = P.stats.find(name);
  if (it != P.stats.end()) return it->second;
  SearchStats* d = nullptr;
  SR_HIP(hipMalloc(&d, sizeof(SearchStats)));
  SR_HIP(hipMemset(d, 0, sizeof(SearchStats)));
  P.stats[name] = d;
  return d;
}

ProfScope::ProfScope(const char* name, hipStream_t s) : stream(s) {
  Profiler& P = profiler();
  if (!P.on.load() || t_arena) return;  // the allocation-free entry point records nothing
  hipEvent_t a = prof_event(), b = prof_event();
  SR_HIP(hipEventRecord(a, s));
  std::lock_guard<std::mutex> lk(P.mu);
  P.recs.push_back({name, a, b});
  slot = int(P.recs.size()) - 1;
}

ProfScope::~ProfScope() {
  if (slot < 0) return;
  Profiler& P = profiler();
  hipEvent_t b;
  {
    std::lock_guard<std::mutex> lk(P.mu);
    b = P.recs[size_t(slot)].b;
  }
  (void)hipEventRecord(b, stream);
}

// -------------------------------------------------------------- options ----

Options& options() {
  static Options o;
  return o;
}

// --------------------------------------------------------- preparation ----

struct Prepared {
  Grid grid;               // points (after the voxel filter) in grid order
  DevBuf<float4> rec;  // grid-order point records (kRec float4 a point: point, normal)
  uint32_t n = 0;
};

struct Receipt {
  uint64_t h2d = 0, d2h = 0, d2d = 0;
  uint32_t stages = 0;
};

// Tuning overrides (environment, read once): SR_GRID_OCCUPANCY = target points
// per occupied grid cell, SR_GICP_PPT = source points per thread in a pass.
static float env_float(const char* name, float dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  char* end = nullptr;
  const float f = std::strtof(v, &end);
  return (end && end != v && f > 0.0f) ? f : dflt;
}

static float knn_occupancy(uint32_t k) {
  static const float forced = env_float("SR_GRID_OCCUPANCY", 0.0f);
  if (forced > 0.0f) return forced;
  return std::max(4.0f, float(k) * 0.35f);
}

// Voxel filter (optional) + grid + plane normals of one cloud. With defer_knn
// the k-NN normals are left to the caller (knn_normals_many over many clouds);
// out.normals is allocated either way.
static void prepare_cloud(const Cloud& c, const sr_gicp_config& cfg, Prepared& out, Receipt& rc,
                          hipStream_t s, bool is_target, float leaf_hint = 0.0f,
                          bool defer_knn = false);

static void prepare(const sr_xyz_view& v, const sr_gicp_config& cfg, Prepared& out, Receipt& rc,
                    hipStream_t s, bool is_target, bool defer_knn = false) {
  if (v.n > 0xFFFFFFF0ull) fail(SR_ERR_INVALID_ARGUMENT, "point clouds above 4G points are not supported");
  Cloud raw;
  pack_cloud(v.x, v.y, v.z, uint32_t(v.n), v.mem == SR_MEM_DEVICE, raw, s, &rc.h2d);
  Cloud* cloud = &raw;
  VoxelOut vox;
  if (cfg.voxel_leaf > 0.0f) {
    voxel_downsample(raw, cfg.voxel_leaf, nullptr, false, vox, s);
    cloud = &vox.centroids;
    rc.stages |= SR_STAGE_VOXEL;
  }
  prepare_cloud(*cloud, cfg, out, rc, s, is_target, cfg.voxel_leaf > 0.0f ? cfg.voxel_leaf : 0.0f,
                defer_knn);
}

// Grid and covariances of a packed device cloud (no voxel filter).
static void prepare_cloud(const Cloud& c, const sr_gicp_config& cfg, Prepared& out, Receipt& rc,
                          hipStream_t s, bool is_target, float leaf_hint, bool defer_knn) {
  const Cloud* cloud = &c;
  out.n = cloud->n;
  uint32_t hb[6];
  host_bounds(*cloud, hb, s);
  const float h = choose_cell_edge(*cloud, hb, knn_occupancy(cfg.k_neighbors), leaf_hint);
  build_grid(*cloud, hb, h, out.grid, s);
  if (is_target) build_empty_radius(out.grid, s);
  rc.stages |= SR_STAGE_GRID;
  out.rec.resize(kRec * std::max<uint32_t>(out.n, 1));
  if (cfg.covariance_radius > 0.0f) {
    Grid rg;
    build_grid(*cloud, hb, cfg.covariance_radius, rg, s);
    DevBuf<float4> orig(std::max<uint32_t>(out.n, 1));
    radius_normals(rg, cfg.covariance_radius, float(cfg.epsilon), orig.p, nullptr, nullptr, s);
    gather_records(out.grid, orig.p, out.rec.p, s);
    SR_HIP(hipStreamSynchronize(s));
  } else if (!defer_knn) {
    knn_normals(out.grid, cfg.k_neighbors, cfg.epsilon, out.rec.p, nullptr, nullptr, s);
  }
  rc.stages |= SR_STAGE_COVARIANCE;
}

static InstConfig inst_config(const sr_gicp_config& c) {
  InstConfig ic{};
  ic.dmax = c.max_correspondence_distance;
  ic.dmax2 = c.max_correspondence_distance * c.max_correspondence_distance;
  ic.one_minus_eps = float(1.0 - c.epsilon);
  ic.max_iter = c.max_iterations;
  ic.min_corr = c.min_correspondences;
  ic.flags = options().match_reuse.load() ? kCfgMatchReuse : 0u;
  ic.trans_eps = c.transformation_epsilon;
  ic.fit_eps = c.fitness_epsilon;
  return ic;
}

static uint32_t pick_ppt(uint64_t points) {
  static const float forced = env_float("SR_GICP_PPT", 0.0f);
  if (forced >= 1.0f) return uint32_t(forced);
  // Enough workgroups to fill 256 CUs several times over before packing more
  // points per thread (fewer block reductions). Chosen from one registration's
  // own size, so a registration sums the same way alone and inside a batch.
  if (points >= (1ull << 23)) return 4;
  if (points >= (1ull << 21)) return 2;
  return 1;
}

// Persistent preprocessing workers. Small clouds cannot fill the GPU one at a
// time and their preparation waits on a few host reads (cell counts, bounds),
// so many clouds are prepared concurrently: each worker thread has its own
// stream and its own allocator free lists, and the threads live as long as
// the process so those free lists keep being reused.
class Workers {
 public:
  static Workers& get() {
    // Eight threads: their host waits overlap each other's device work (measured
    // against 4 and 16 on 256 small pairs); SR_PREP_WORKERS overrides.
    static Workers* w =
        new Workers(unsigned(std::min(64.0f, std::max(1.0f, env_float("SR_PREP_WORKERS", 8.0f)))));
    return *w;
  }
  // Runs fn(i, stream) for i in [0, n) on the workers, each worker's stream
  // synchronised before run returns; rethrows the first error.
  void run(size_t n, const std::function<void(size_t, hipStream_t)>& fn) {
    if (n == 0) return;
    std::unique_lock<std::mutex> call(call_mu_);  // one batch job at a time
    std::unique_lock<std::mutex> lk(mu_);
    job_ = &fn;
    n_ = n;
    next_ = 0;
    busy_ = nthreads_;
    err_ = nullptr;
    ++gen_;
    cv_.notify_all();
    done_cv_.wait(lk, [&] { return busy_ == 0; });
    job_ = nullptr;
    if (err_) std::rethrow_exception(err_);
  }

 private:
  explicit Workers(unsigned n) : nthreads_(n) {
    for (unsigned i = 0; i < n; ++i) std::thread([this] { loop(); }).detach();
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(size_t, hipStream_t)>* job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        job = job_;
      }
      std::exception_ptr err;
      try {
        hipStream_t s = thread_stream();
        for (;;) {
          size_t i;
          {
            std::lock_guard<std::mutex> lk(mu_);
            if (next_ >= n_ || err_) break;
            i = next_++;
          }
          (*job)(i, s);
        }
        SR_HIP(hipStreamSynchronize(s));
      } catch (...) {
        err = std::current_exception();
      }
      std::lock_guard<std::mutex> lk(mu_);
      if (err && !err_) err_ = err;
      if (--busy_ == 0) done_cv_.notify_all();
    }
  }
  const unsigned nthreads_;
  std::mutex call_mu_, mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(size_t, hipStream_t)>* job_ = nullptr;
  size_t n_ = 0, next_ = 0;
  unsigned busy_ = 0;
  uint64_t gen_ = 0;
  std::exception_ptr err_;
};

// Prepares many clouds on the workers; out[i] receives views[i] prepared.
static void prepare_many(const std::vector<const sr_xyz_view*>& views,
                         const std::vector<sr_gicp_config>& cfgs, std::vector<Prepared>& out,
                         bool is_target, Receipt& rc) {
  std::mutex mu;
  Workers::get().run(views.size(), [&](size_t i, hipStream_t s) {
    Receipt local;
    prepare(*views[i], cfgs[i], out[i], local, s, is_target);
    std::lock_guard<std::mutex> lk(mu);
    rc.h2d += local.h2d;
    rc.d2h += local.d2h;
    rc.d2d += local.d2d;
    rc.stages |= local.stages;
  });
}

// Prepares the targets and sources of many pairs: voxel filters and grids in
// one job on the workers (so they never idle between the two lists), then the
// k-NN normals of all clouds in one launch per k on the caller's stream. Small
// clouds each leave most of the chip idle while their slowest points finish;
// in one launch those tails overlap.
// Flat buffers the prepared clouds of a segmented preparation view into.
struct SegHolder {
  GridSet set;
  DevBuf<float4> rec;
};

// Preparation of many device-resident clouds of one configuration (voxel leaf
// set, k-NN covariances) as segments of flat buffers: each stage is one set of
// launches over all clouds and waits on the host at most a few times in all,
// instead of a few dozen launches and four waits per cloud. Cloud c's results
// are those of prepare() on it alone. Returns null when the clouds do not
// qualify (then nothing was done).
static std::shared_ptr<SegHolder> prepare_segmented(const std::vector<const sr_xyz_view*>& views,
                                                    const std::vector<bool>& is_target,
                                                    const sr_gicp_config& cfg,
                                                    std::vector<Prepared*>& out, Receipt& rc,
                                                    hipStream_t s) {
  if (!(cfg.voxel_leaf > 0.0f) || cfg.covariance_radius > 0.0f || views.empty()) return nullptr;
  uint64_t total = 0;
  for (const sr_xyz_view* v : views) {
    if (v->mem != SR_MEM_DEVICE || v->n == 0) return nullptr;
    total += v->n;
  }
  if (total > 0xFFFFFFF0ull) return nullptr;
  const uint32_t nseg = uint32_t(views.size());
  std::vector<SegSource> src(nseg);
  std::vector<uint32_t> off(nseg + 1, 0);
  for (uint32_t c = 0; c < nseg; ++c) {
    src[c] = SegSource{views[c]->x, views[c]->y, views[c]->z, uint32_t(views[c]->n), off[c]};
    off[c + 1] = off[c] + uint32_t(views[c]->n);
  }
  Cloud raw;
  DevBuf<uint8_t> pack_tables;
  const uint32_t* raw_bounds = nullptr;
  pack_many(src, raw, s, &pack_tables, &raw_bounds);  // also the segments' bounds
  VoxelOut vox;
  std::vector<uint32_t> voff;
  voxel_downsample_many(raw, off, cfg.voxel_leaf, vox, voff, s, raw_bounds);
  rc.stages |= SR_STAGE_VOXEL;
  raw.pts.release();
  std::vector<uint32_t> hb(6 * size_t(nseg));
  {
    DevBuf<uint32_t> b(6 * size_t(nseg));
    seg_bounds(vox.centroids.pts.p, voff, b.p, s);
    download(hb.data(), b.p, sizeof(uint32_t) * hb.size(), s);
  }
  std::vector<float> h;
  choose_cell_edges_many(vox.centroids, voff, hb, knn_occupancy(cfg.k_neighbors), cfg.voxel_leaf,
                         h, s);
  auto hold = std::make_shared<SegHolder>();
  // the empty radius serves the targets' searches: computed for the leading
  // run of target segments (all of them when the targets come first)
  uint32_t empty_segs = 0;
  while (empty_segs < nseg && is_target[empty_segs]) ++empty_segs;
  for (uint32_t c = empty_segs; c < nseg; ++c)
    if (is_target[c]) empty_segs = nseg;
  build_grids_many(vox.centroids, voff, hb, h, empty_segs, hold->set, s);
  rc.stages |= SR_STAGE_GRID;
  const GridSet& G = hold->set;
  hold->rec.resize(kRec * std::max<uint32_t>(voff.back(), 1));
  std::vector<KnnCloud> kc(nseg);
  for (uint32_t c = 0; c < nseg; ++c) {
    const uint32_t n = voff[c + 1] - voff[c];
    Prepared& P = *out[c];
    P.n = n;
    P.grid.desc = G.desc[c];
    P.grid.cell_start = DevBuf<uint32_t>::view(G.cell_start.p + G.cs_off[c], G.desc[c].ncells + 1);
    P.grid.pts = DevBuf<float4>::view(G.pts.p + voff[c], n);
    if (is_target[c])
      P.grid.empty_radius = DevBuf<uint8_t>::view(G.empty_radius.p + G.cell_off[c], G.desc[c].ncells);
    P.rec = DevBuf<float4>::view(hold->rec.p + kRec * size_t(voff[c]), kRec * size_t(n));
    kc[c] = knn_cloud(P.grid, P.rec.p, nullptr, nullptr);
  }
  knn_normals_many(kc, cfg.k_neighbors, cfg.epsilon, s);
  rc.stages |= SR_STAGE_COVARIANCE;
  return hold;
}

static std::shared_ptr<SegHolder> prepare_pairs(const std::vector<const sr_xyz_view*>& tviews,
                                                const std::vector<const sr_xyz_view*>& sviews,
                                                const std::vector<sr_gicp_config>& cfgs,
                                                std::vector<Prepared>& tout,
                                                std::vector<Prepared>& sout, Receipt& rc) {
  {
    bool same = true;
    for (const sr_gicp_config& c : cfgs) same = same && std::memcmp(&c, &cfgs[0], sizeof(c)) == 0;
    if (same && !cfgs.empty()) {
      std::vector<const sr_xyz_view*> views;
      std::vector<bool> tgt;
      std::vector<Prepared*> outs;
      for (size_t i = 0; i < tviews.size(); ++i) {
        views.push_back(tviews[i]);
        tgt.push_back(true);
        outs.push_back(&tout[i]);
      }
      for (size_t i = 0; i < sviews.size(); ++i) {
        views.push_back(sviews[i]);
        tgt.push_back(false);
        outs.push_back(&sout[i]);
      }
      auto hold = prepare_segmented(views, tgt, cfgs[0], outs, rc, thread_stream());
      if (hold) return hold;
    }
  }
  std::mutex mu;
  Workers::get().run(2 * tviews.size(), [&](size_t j, hipStream_t s) {
    const size_t i = j / 2;
    Receipt local;
    const bool defer = !(cfgs[i].covariance_radius > 0.0f);
    if (j % 2 == 0)
      prepare(*tviews[i], cfgs[i], tout[i], local, s, true, defer);
    else
      prepare(*sviews[i], cfgs[i], sout[i], local, s, false, defer);
    std::lock_guard<std::mutex> lk(mu);
    rc.h2d += local.h2d;
    rc.d2h += local.d2h;
    rc.d2d += local.d2d;
    rc.stages |= local.stages;
  });
  std::map<uint32_t, std::pair<double, std::vector<KnnCloud>>> by_k;
  for (size_t i = 0; i < tviews.size(); ++i) {
    if (cfgs[i].covariance_radius > 0.0f) continue;
    auto& e = by_k[cfgs[i].k_neighbors];
    if (e.second.empty()) e.first = cfgs[i].epsilon;
    if (e.first != cfgs[i].epsilon) fail(SR_ERR_INVALID_ARGUMENT, "pairs of one batch must share epsilon");
    e.second.push_back(knn_cloud(tout[i].grid, tout[i].rec.p, nullptr, nullptr));
    e.second.push_back(knn_cloud(sout[i].grid, sout[i].rec.p, nullptr, nullptr));
  }
  hipStream_t s = thread_stream();
  for (auto& kv : by_k) knn_normals_many(kv.second.second, kv.first, kv.second.first, s);
  if (!by_k.empty()) rc.stages |= SR_STAGE_COVARIANCE;
  return nullptr;
}

// Runs the Gauss-Newton loop for `count` (source, target) pairs in one set of
// launches and fills `out`. Each registration keeps its own blocks, partials and
// state, so results do not depend on the batch.
// Launch descriptors live in per-thread buffers that only ever grow, so the
// kernel arguments of a batch of a given shape stay the same from call to call
// and its captured pass loop (batch_run) can be replayed.
struct LoopArgs {
  DevBuf<IterInstance> insts;
  DevBuf<uint32_t> block_inst;
  DevBuf<HardQuery> queue;  // pass scratch (PassScratch)
  DevBuf<uint32_t> ctl, done;
};

// Grows a pass scratch to `points` queue slots and `blocks` flags and masks
// (the masks, four u64 words a block, follow the flags in the same buffer);
// the queue length starts at zero (the finalize kernel keeps it so between
// passes).
static PassScratch grow_scratch(DevBuf<HardQuery>& queue, DevBuf<uint32_t>& ctl,
                                DevBuf<uint32_t>& done, uint64_t points, uint32_t blocks,
                                hipStream_t s) {
  queue.resize(std::max<uint64_t>(points, 1));
  const uint64_t flags = (uint64_t(std::max<uint32_t>(blocks, 1)) + 1) & ~1ull;  // 8-byte aligned
  done.resize(flags + 8 * uint64_t(std::max<uint32_t>(blocks, 1)));
  ctl.resize(1);
  fill_u32(ctl.p, 0u, 1, s);
  return PassScratch{queue.p, ctl.p, done.p,
                     reinterpret_cast<unsigned long long*>(done.p + flags)};
}
static LoopArgs& thread_loop_args() {
  static thread_local LoopArgs* a = new LoopArgs();  // lives until the thread ends
  return *a;
}

struct Batch {
  // Launch descriptors and pass scratch: the thread's persistent buffers, or,
  // inside the allocation-free entry point, buffers of this call (from its
  // arena, released with the batch).
  std::unique_ptr<LoopArgs> own = arena_active() ? std::make_unique<LoopArgs>() : nullptr;
  LoopArgs& la = own ? *own : thread_loop_args();
  DevBuf<IterInstance>& insts = la.insts;
  DevBuf<uint32_t>& block_inst = la.block_inst;
  DevBuf<InstState> states;
  DevBuf<float> partials;
  DevBuf<double> init;
  DevBuf<uint32_t> match;
  DevBuf<float> budget;
  DevBuf<float4> seen;
  uint32_t total_blocks = 0, count = 0, max_passes = 0;
  uint64_t total_pts = 0;
  PassScratch scratch{};
  uint32_t hard_blocks = 1;
};

// Work split of one registration's source over pass blocks.
static uint32_t instance_blocks(uint32_t n_src) {
  const uint32_t ppt = pick_ppt(n_src);
  return std::max<uint32_t>(1, (n_src + kBlock * ppt - 1) / (kBlock * ppt));
}

// Launch descriptor of one registration: its clouds, config and the device
// state, partials (instance_blocks x 64 floats) and match (n_src words) it owns.
static IterInstance describe_instance(const Prepared& src, const Prepared& tgt,
                                      const sr_gicp_config& cfg, InstState* state, float* partials,
                                      uint32_t* match, float* budget, float4* seen,
                                      uint32_t block_begin) {
  IterInstance I{};
  I.src_rec = src.rec.p;
  I.tgt = tgt.grid.pts.p;
  I.tgt_rec = tgt.rec.p;
  I.cell_start = tgt.grid.cell_start.p;
  I.empty_radius = tgt.grid.empty_radius.p;
  I.match = match;
  I.budget = budget;
  I.seen = seen;
  {
    // the source grid's box holds every source point
    const GridDesc& d = src.grid.desc;
    const float lo[3] = {d.ox, d.oy, d.oz};
    const float ext[3] = {d.h * float(d.dx), d.h * float(d.dy), d.h * float(d.dz)};
    double r2 = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double m = std::max(std::fabs(double(lo[a])), std::fabs(double(lo[a]) + ext[a]));
      r2 += m * m;
    }
    I.src_radius = float(std::sqrt(r2) * (1.0 + 1e-6) + 1e-6);
  }
  I.grid = tgt.grid.desc;
  I.cfg = inst_config(cfg);
  I.state = state;
  I.partials = partials;
  I.n_src = src.n;
  I.block_begin = block_begin;
  I.nblocks = instance_blocks(src.n);
  I.ppt = pick_ppt(src.n);
  return I;
}

static void batch_setup(const std::vector<const Prepared*>& src, const std::vector<const Prepared*>& tgt,
                        const std::vector<sr_gicp_config>& cfgs, Batch& B, hipStream_t s,
                        const double* pose7 = nullptr) {
  const uint32_t count = uint32_t(src.size());
  B.count = count;
  uint64_t total_pts = 0;
  for (auto* p : src) total_pts += p->n;
  std::vector<IterInstance> hi(count);
  std::vector<uint32_t> hbi;
  std::vector<double> init(size_t(7) * count);
  uint32_t blocks = 0;
  B.max_passes = 0;
  std::vector<uint32_t> nbl(count);
  for (uint32_t i = 0; i < count; ++i) {
    nbl[i] = instance_blocks(src[i]->n);
    blocks += nbl[i];
  }
  B.states.resize(count);
  B.partials.resize(size_t(blocks) * 64);
  B.match.resize(std::max<uint64_t>(total_pts, 1));
  B.budget.resize(std::max<uint64_t>(total_pts, 1));
  B.seen.resize(std::max<uint64_t>(total_pts, 1));
  uint32_t begin = 0;
  uint64_t src_begin = 0;
  for (uint32_t i = 0; i < count; ++i) {
    hi[i] = describe_instance(*src[i], *tgt[i], cfgs[i], B.states.p + i,
                              B.partials.p + size_t(begin) * 64, B.match.p + src_begin,
                              B.budget.p + src_begin, B.seen.p + src_begin, begin);
    src_begin += src[i]->n;
    for (uint32_t b = 0; b < nbl[i]; ++b) hbi.push_back(i);
    begin += nbl[i];
    for (int a = 0; a < 4; ++a) init[7 * size_t(i) + a] = double(cfgs[i].init_q[a]);
    for (int a = 0; a < 3; ++a) init[7 * size_t(i) + 4 + a] = double(cfgs[i].init_t[a]);
    if (pose7)
      for (int a = 0; a < 7; ++a) init[7 * size_t(i) + a] = pose7[7 * size_t(i) + a];
    B.max_passes = std::max(B.max_passes, cfgs[i].max_iterations + 1);
  }
  B.total_blocks = blocks;
  B.total_pts = total_pts;
  {
    LoopArgs& la = B.la;
    B.scratch = grow_scratch(la.queue, la.ctl, la.done, total_pts, blocks, s);
    B.hard_blocks = hard_stage_blocks(total_pts);
  }
  B.insts.resize(count);
  B.block_inst.resize(blocks);
  B.init.resize(init.size());
  upload(B.insts.p, hi.data(), sizeof(IterInstance) * count, s);
  upload(B.block_inst.p, hbi.data(), sizeof(uint32_t) * blocks, s);
  upload(B.init.p, init.data(), sizeof(double) * init.size(), s);
  reset_batch_state(B.states.p, B.init.p, count, B.match.p, B.budget.p, B.seen.p, total_pts, s);
}

// ---- launch tuner -------------------------------------------------------------
//
// The pass loop's launch choices (PassTuning: lanes per queued point in the
// ring search, that stage's workgroup cap, the size from which the first stage
// splits into two kernels) change no result: matches, budgets and sums are the
// same whichever is used (the match-reuse and two-kernel tests show the same
// for the choices they cover). Their speed depends on the problem, so the
// first pass loop of each problem class (device architecture; one or many
// registrations; log2 of the source points) runs once per candidate from the
// same initial state, timed with HIP events, and the fastest is kept for the
// class. Loops too small for the choice to matter (one registration under
// 128k source points) use the defaults. SR_TUNE_CACHE names a file the decisions
// are read from and appended to, so a process can start tuned.

struct TuneEntry {
  PassTuning best;
  std::vector<float> ms;  // per candidate
};

static std::vector<PassTuning> tune_candidates(bool batch) {
  std::vector<PassTuning> c;
  auto add = [&](int g, uint32_t cap, uint32_t split) { c.push_back(PassTuning{g, cap, split}); };
  add(8, 4096, 4096);  // the defaults first
  if (batch) add(4, 4096, 4096);  // single registrations are slower with 4 lanes
  add(16, 4096, 4096);
  add(8, 2048, 4096);
  add(8, 8192, 4096);
  if (batch) {
    add(8, 4096, 1u << 30);  // never split
    add(8, 4096, 16384);
  }
  return c;
}

struct Tuner {
  std::mutex mu;
  std::map<std::pair<std::string, uint32_t>, TuneEntry> done;
  bool loaded = false;
};
static Tuner& tuner() {
  static Tuner* t = new Tuner();
  return *t;
}

static std::string device_arch() {
  static std::string arch = [] {
    hipDeviceProp_t p;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return std::string("unknown");
    std::string a = p.gcnArchName;
    return a.substr(0, a.find(':'));
  }();
  return arch;
}

static uint32_t tune_class(const Batch& B) {
  uint32_t l = 0;
  while ((uint64_t(1) << (l + 1)) <= B.total_pts) ++l;
  return (B.count > 1 ? 0x100u : 0u) | l;
}

static void tuner_load(Tuner& T) {
  if (T.loaded) return;
  T.loaded = true;
  const char* path = std::getenv("SR_TUNE_CACHE");
  if (!path || !*path) return;
  std::ifstream f(path);
  std::string arch;
  uint32_t cls, cap, split;
  int group;
  while (f >> arch >> cls >> group >> cap >> split) T.done[{arch, cls}].best = PassTuning{group, cap, split};
}

static void tuner_store(const std::string& arch, uint32_t cls, const PassTuning& t) {
  const char* path = std::getenv("SR_TUNE_CACHE");
  if (!path || !*path) return;
  std::ofstream f(path, std::ios::app);
  f << arch << ' ' << cls << ' ' << t.group << ' ' << t.hard_cap << ' ' << t.split_blocks << '\n';
}

static void run_passes(Batch& B, hipStream_t s, const PassTuning& t) {
  const uint32_t hb = hard_stage_blocks(B.total_pts, t.hard_cap);
  for (uint32_t k = 0; k < B.max_passes; ++k)
    gicp_pass(B.insts.p, B.block_inst.p, B.total_blocks, B.count, B.scratch, hb, s, true, t);
}

// The registrations' starting state (as batch_setup leaves it).
static void reset_batch(Batch& B, hipStream_t s) {
  reset_batch_state(B.states.p, B.init.p, B.count, B.match.p, B.budget.p, B.seen.p, B.total_pts, s);
}

// Times every candidate on this batch (eager launches) and returns the
// fastest; the batch is left as the last run finished it, which is the result
// every candidate gives.
static PassTuning tune_batch(Batch& B, hipStream_t s, TuneEntry& e) {
  const std::vector<PassTuning> cand = tune_candidates(B.count > 1);
  hipEvent_t a, b;
  SR_HIP(hipEventCreate(&a));
  SR_HIP(hipEventCreate(&b));
  // Two rounds over the candidates, each keeping its faster run: the first
  // loop of a process runs on cold caches and allocator, which alone would
  // penalise whichever candidate went first.
  e.ms.assign(cand.size(), INFINITY);
  size_t best = 0;
  for (int round = 0; round < 2; ++round) {
    for (size_t c = 0; c < cand.size(); ++c) {
      if (round > 0 || c > 0) reset_batch(B, s);
      float ms = 0.0f;
      SR_HIP(hipEventRecord(a, s));
      run_passes(B, s, cand[c]);
      SR_HIP(hipEventRecord(b, s));
      SR_HIP(hipEventSynchronize(b));
      SR_HIP(hipEventElapsedTime(&ms, a, b));
      e.ms[c] = std::min(e.ms[c], ms);
      if (e.ms[c] < e.ms[best]) best = c;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  e.best = cand[best];
  return e.best;
}

// The launch choices for this batch's pass loop (tuning the class first if it
// is new; `tuned` reports whether the loop has already run doing so).
static PassTuning choose_tuning(Batch& B, hipStream_t s, bool eager_ok, bool& tuned) {
  tuned = false;
  const int64_t forced = options().tune_force.load();
  if (forced >= 0) {
    const std::vector<PassTuning> cand = tune_candidates(B.count > 1);
    return cand[size_t(forced) % cand.size()];
  }
  if (!options().autotune.load()) return PassTuning{};
  if (B.count == 1 && B.total_pts < (1u << 17)) return PassTuning{};
  Tuner& T = tuner();
  const std::pair<std::string, uint32_t> key{device_arch(), tune_class(B)};
  {
    std::lock_guard<std::mutex> lk(T.mu);
    tuner_load(T);
    auto it = T.done.find(key);
    if (it != T.done.end()) return it->second.best;
  }
  if (!eager_ok) return PassTuning{};
  TuneEntry e;
  const PassTuning best = tune_batch(B, s, e);
  tuned = true;
  std::lock_guard<std::mutex> lk(T.mu);
  if (T.done.emplace(key, e).second) tuner_store(key.first, key.second, best);
  return best;
}

// The Gauss-Newton loop of a batch: max_passes pairs of (pass, finalize)
// launches. Outside profiling the loop is captured once per launch shape as a
// HIP graph and replayed: its kernel arguments are the per-thread descriptor
// buffers and sizes, while everything a registration changes (clouds, states,
// partials) is read through the descriptors at run time. The cache is per host
// thread, since a graph replays on that thread's stream, and every entry point
// synchronises before it returns, so no cached graph is in flight when one is
// evicted.
struct GraphEntry {
  const void* insts;
  const void* block_inst;
  const void* queue;
  const void* done;
  uint32_t blocks, count, passes, hard_blocks;
  int group;
  uint32_t hard_cap, split;
  hipGraphExec_t exec;
};

static void batch_run(Batch& B, hipStream_t s) {
  // eager when profiling (so each launch can be timed), inside the
  // allocation-free entry point (instantiating a graph allocates) and with
  // the pass_graphs option off
  const bool eager = profiler().on.load() != 0 || arena_active() || options().pass_graphs.load() == 0;
  bool tuned = false;
  const PassTuning tune = choose_tuning(B, s, !eager, tuned);
  if (tuned) return;  // the tuning runs finished the loop
  B.hard_blocks = hard_stage_blocks(B.total_pts, tune.hard_cap);
  if (eager) {
    run_passes(B, s, tune);
    return;
  }
  static thread_local std::vector<GraphEntry>* cache = new std::vector<GraphEntry>();
  // Shapes seen once: capturing and instantiating a graph costs more than the
  // eager launches of one loop, so a shape is captured on its second run (a
  // stream of one-off shapes, such as odometry scans, stays eager).
  static thread_local std::vector<GraphEntry>* seen = new std::vector<GraphEntry>();
  const GraphEntry key{B.insts.p, B.block_inst.p, B.scratch.queue, B.scratch.done, B.total_blocks,
                       B.count, B.max_passes, B.hard_blocks, tune.group, tune.hard_cap,
                       tune.split_blocks, nullptr};
  auto same = [&](const GraphEntry& e) {
    return e.insts == key.insts && e.block_inst == key.block_inst && e.queue == key.queue &&
           e.done == key.done && e.blocks == key.blocks && e.count == key.count &&
           e.passes == key.passes && e.hard_blocks == key.hard_blocks && e.group == key.group &&
           e.hard_cap == key.hard_cap && e.split == key.split;
  };
  hipGraphExec_t exec = nullptr;
  for (const GraphEntry& e : *cache)
    if (same(e)) exec = e.exec;
  if (!exec) {
    bool again = false;
    for (const GraphEntry& e : *seen) again = again || same(e);
    if (!again) {
      if (seen->size() >= 16) seen->erase(seen->begin());
      seen->push_back(key);
      run_passes(B, s, tune);
      return;
    }
    SR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipGraph_t graph = nullptr;
    try {
      run_passes(B, s, tune);
    } catch (...) {
      (void)hipStreamEndCapture(s, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    SR_HIP(hipStreamEndCapture(s, &graph));
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SR_HIP(e);
    if (cache->size() >= 8) {
      SR_HIP(hipStreamSynchronize(s));
      (void)hipGraphExecDestroy(cache->front().exec);
      cache->erase(cache->begin());
    }
    GraphEntry ent = key;
    ent.exec = exec;
    cache->push_back(ent);
  }
  SR_HIP(hipGraphLaunch(exec, s));
}

static int32_t state_to_result(const InstState& st, const sr_gicp_config& cfg, const Prepared& src,
                               const Prepared& tgt, sr_result* out, std::string* msg) {
  sr_result r;
  std::memset(&r, 0, sizeof(r));
  for (int a = 0; a < 4; ++a) {
    r.q64[a] = st.q[a];
    r.q[a] = float(st.q[a]);
  }
  for (int a = 0; a < 3; ++a) {
    r.t64[a] = st.t[a];
    r.t[a] = float(st.t[a]);
  }
  r.fitness = st.fitness;
  r.iterations = st.iter;
  r.converged = uint8_t(st.converged);
  r.source_points = src.n;
  r.target_points = tgt.n;
  r.correspondences = st.count;
  if (out) *out = r;
  if (st.status == 1) {
    *msg = "GICP found only " + std::to_string(st.count) + " correspondences, minimum is " +
           std::to_string(cfg.min_correspondences);
    return SR_ERR_INVALID_ARGUMENT;
  }
  if (st.status == 2) {
    *msg = "GICP normal equations were singular";
    return SR_ERR_INVALID_ARGUMENT;
  }
  return SR_OK;
}

// k_neighbors bounds: the reference's lower bound (gicp.rs:113-117) and this
// backend's upper bound on the k-NN covariance networks (the radius path does
// not use k).
static void check_k(uint32_t k, float covariance_radius) {
  if (k < 3) fail(SR_ERR_INVALID_ARGUMENT, "GICP needs k_neighbors >= 3 for covariance estimation");
  if (k > kMaxKnnK && !(covariance_radius > 0.0f))
    fail(SR_ERR_INVALID_ARGUMENT, "GICP k-NN covariances on the HIP backend support k_neighbors <= " +
                                      std::to_string(kMaxKnnK));
}

// A request against a resident map must describe the map's target the way the
// map was built: the target normals were computed with the map's k_neighbors
// / covariance_radius and voxel leaf, and the combined covariance uses one
// epsilon for both clouds (gicp.rs:119-123 computes both with one config).
static void check_map_config(const sr_gicp_config& map_cfg, const sr_gicp_config& cfg) {
  const bool radius = map_cfg.covariance_radius > 0.0f, r2 = cfg.covariance_radius > 0.0f;
  if (map_cfg.epsilon != cfg.epsilon || map_cfg.voxel_leaf != cfg.voxel_leaf || radius != r2 ||
      (radius && map_cfg.covariance_radius != cfg.covariance_radius) ||
      (!radius && map_cfg.k_neighbors != cfg.k_neighbors))
    fail(SR_ERR_INVALID_ARGUMENT,
         "config must match the map's epsilon, voxel_leaf, k_neighbors and covariance_radius");
}

static void validate_cfg(const sr_xyz_view* source, const sr_xyz_view* target,
                         const sr_gicp_config* cfg) {
  if (!cfg) fail(SR_ERR_INVALID_ARGUMENT, "config must not be null");
  if ((source && source->n == 0) || (target && target->n == 0) || !source || !target)
    fail(SR_ERR_INVALID_ARGUMENT, "GICP requires non-empty source and target point clouds");
  check_k(cfg->k_neighbors, cfg->covariance_radius);
  if (!(cfg->voxel_leaf <= 0.0f) && !(cfg->voxel_leaf > 0.0f))
    fail(SR_ERR_INVALID_ARGUMENT, "leaf_size must be greater than zero");
}

static void validate_view(const sr_xyz_view* v) {
  if (!v) fail(SR_ERR_INVALID_ARGUMENT, "point cloud view must not be null");
  if (v->n > 0 && (!v->x || !v->y || !v->z))
    fail(SR_ERR_INVALID_ARGUMENT, "point cloud view has null columns");
  if (v->mem != SR_MEM_HOST && v->mem != SR_MEM_DEVICE)
    fail(SR_ERR_INVALID_ARGUMENT, "point cloud view mem must be 0 (host) or 1 (device)");
}

// Resolves an ExecutionPolicy for this backend (execution.rs:17-58 contract):
// Gpu(Cpu) is invalid; Gpu(Hip) is strict; Auto uses HIP when it is usable and
// otherwise reports UNAVAILABLE so the caller runs its own CPU path; CPU
// policies are the caller's CPU implementation.
static sr_policy resolve(sr_policy p) {
  if (p.kind == SR_GPU && p.device == SR_DEV_CPU)
    fail(SR_ERR_INVALID_ARGUMENT, "GPU execution policy cannot target the CPU device");
  if (p.kind == SR_CPU_SINGLE || p.kind == SR_CPU_PARALLEL)
    fail(SR_STATUS_UNAVAILABLE, "CPU policies run in the caller's CPU implementation; this library is the HIP backend");
  if (p.kind == SR_GPU && p.device != SR_DEV_HIP)
    fail(SR_ERR_INVALID_ARGUMENT, "unsupported GPU device kind for this backend (expected Hip)");
  if (p.kind != SR_GPU && p.kind != SR_AUTO) fail(SR_ERR_INVALID_ARGUMENT, "unknown execution policy");
  try {
    ensure_device();
  } catch (const Error& e) {
    if (p.kind == SR_AUTO) throw;  // UNAVAILABLE: Auto may fall back
    fail(SR_ERR_INVALID_ARGUMENT, e.msg);
  }
  return sr_policy{SR_GPU, SR_DEV_HIP};
}

}  // namespace sr

// ------------------------------------------------- allocation-free entry ----
//
// sr_gicp_align_noalloc runs sr_gicp_align's kernels with every device buffer
// carved from `scratch`, a file-scope device array each GPU holds for the
// life of the process, so the call makes no hipMalloc / hipFree and no
// allocating library call (and captures no graph, since instantiating one
// allocates). Calls on one device are serialised by a lock and each ends with
// its stream drained, so they never share the array's contents at the same
// time; the array keeps the last call's data, and every kernel of the path
// clears or overwrites what it reads (as with reused allocator blocks).

namespace sr {
constexpr size_t kScratchBytes = size_t(256) << 20;  // per device
constexpr uint64_t kNoAllocMaxPoints = 2000000;     // per cloud
constexpr int kMaxDevices = 64;
}  // namespace sr

static __device__ __align__(256) unsigned char scratch[sr::kScratchBytes];

namespace sr {

struct ScratchDevice {
  std::mutex call;      // one allocation-free call on the device at a time
  char* ptr = nullptr;  // device address of `scratch` on this device
  Arena arena;
  size_t last_peak = 0;
};

static ScratchDevice& scratch_device(int dev) {
  static ScratchDevice devs[kMaxDevices];
  static std::mutex setup;
  if (dev < 0 || dev >= kMaxDevices) fail(SR_ERR_INVALID_ARGUMENT, "device id out of range");
  ScratchDevice& d = devs[dev];
  std::lock_guard<std::mutex> lk(setup);
  if (!d.ptr) {
    // Once per device, with that device current: the host name `scratch` is
    // only a stand-in; kernels and copies take this address.
    void* p = nullptr;
    SR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(SR_ERR_INVALID_ARGUMENT, "scratch array has no device address");
    d.ptr = static_cast<char*>(p);
  }
  return d;
}

// Installs the device's arena on this thread for one call; on exit (normal or
// error) the stream is drained before the lock is released.
struct ArenaScope {
  ScratchDevice& d;
  hipStream_t s;
  ArenaScope(ScratchDevice& dev, hipStream_t st) : d(dev), s(st) {
    d.arena.reset(d.ptr, kScratchBytes);
    t_arena = &d.arena;
  }
  ~ArenaScope() {
    (void)hipStreamSynchronize(s);
    d.last_peak = d.arena.peak;
    t_arena = nullptr;
  }
};

}  // namespace sr

// ------------------------------------------------------------------ ABI ----

using namespace sr;

struct sr_map {
  Prepared tgt;
  sr_gicp_config cfg;
};

// ---- continuous-batching server ------------------------------------------
//
// Requests queue on the host; each scheduling step (server_step) runs one
// Gauss-Newton pass (+ finalize) over every active request in a single launch
// set, while the requests admitted in the same step are prepared on a worker
// stream, so the two overlap on the device. A request owns its prepared
// source, state, partials and match buffer from admission to retirement, and
// its sums never depend on who else is active: its result is the one
// sr_map_align returns for it alone.
//
// The host does not wait for each pass: the states come back through an
// asynchronous copy into pinned memory after every step, and a step retires
// the requests a completed snapshot shows finished. At most kServerDepth steps
// are in flight; a step waits only for the one kServerDepth steps back.

namespace sr {

constexpr int kServerDepth = 2;
// Upper bound of the transient device bytes per raw point while a request is
// prepared: host columns staged on the device (12), the packed cloud (16), the
// voxel table (up to 4 slots of 16 B per point), its per-point work arrays
// (9 x 4) and the centroids (16).
constexpr double kPrepBytesPerPoint = 12 + 16 + 64 + 36 + 16;

struct ServerRequest {
  uint64_t ticket = 0;
  const sr_map* map = nullptr;
  std::vector<float> x, y, z;
  uint64_t reserve = 0;  // device bytes charged against the budget
  std::chrono::steady_clock::time_point submitted;
  Prepared src;
  std::shared_ptr<SegHolder> hold;  // flat buffers the prepared source views into
  uint32_t slot = 0;  // index of its state in the server's state array
  DevBuf<float> partials;
  DevBuf<uint32_t> match;
  DevBuf<float> budget;
  DevBuf<float4> seen;
};

struct ServerDone {
  int32_t code = 0;
  std::string msg;
  sr_result result{};
  double latency_s = 0.0;
};

// One in-flight step's state snapshot: the states copied after its launches,
// the ticket each slot held then, and the event that marks the copy done.
struct ServerSnap {
  InstState* states = nullptr;  // pinned host memory, one per slot
  std::vector<uint64_t> ticket;  // per slot (0: free)
  hipEvent_t done = nullptr;
  bool pending = false;
};

}  // namespace sr

struct sr_server {
  sr_server_config cfg{};
  sr_gicp_config gicp{};
  std::mutex mu;  // queue, done, finished, next_ticket (submit/poll may run beside a step)
  std::deque<std::unique_ptr<sr::ServerRequest>> queue;
  std::vector<std::unique_ptr<sr::ServerRequest>> active;
  std::unordered_map<uint64_t, sr::ServerDone> done;
  std::deque<uint64_t> finished;  // finished tickets not yet listed by sr_server_finished
  uint64_t next_ticket = 1;
  uint64_t bytes_active = 0;
  double bytes_per_point = 64.0;  // admission estimate per raw source point (learned)
  sr::DevBuf<sr::InstState> states;  // one slot per possible active request
  std::vector<uint32_t> free_slots;
  sr::DevBuf<sr::IterInstance> insts;
  sr::DevBuf<uint32_t> block_inst;
  sr::DevBuf<double> init;
  sr::DevBuf<sr::HardQuery> pass_queue;  // pass scratch
  sr::DevBuf<uint32_t> pass_ctl, pass_done;
  sr::ServerSnap snaps[sr::kServerDepth];
  uint64_t steps = 0;
  ~sr_server() {
    for (auto& sn : snaps) {
      if (sn.states) (void)hipHostFree(sn.states);
      if (sn.done) (void)hipEventDestroy(sn.done);
    }
  }
};

namespace sr {

// Per-request GICP config: the server's iteration cap and Delta-pose threshold
// on top of the shared GICP settings.
static sr_gicp_config request_config(const sr_server_config& sc, const sr_gicp_config& g) {
  sr_gicp_config c = g;
  if (sc.max_iterations) c.max_iterations = sc.max_iterations;
  if (sc.delta_pose_epsilon > 0.0) c.transformation_epsilon = sc.delta_pose_epsilon;
  return c;
}

// Device bytes a prepared request holds while active: its grid (cell starts,
// points, order, bitmaps), normals, match, budget and last-search point words,
// block partials, and its share of the pass queue.
static uint64_t resident_bytes(const Prepared& p) {
  const uint64_t n = p.n;
  const GridDesc& d = p.grid.desc;
  const uint64_t words = uint64_t(d.wpr) * uint64_t(d.dy) * uint64_t(d.dz) +
                         uint64_t(d.wps) * uint64_t(d.dz);
  return 4 * (uint64_t(d.ncells) + 1) + n * (16 + 4 + 16 + 4 + 4 + 16 + sizeof(HardQuery)) + 4 * words +
         uint64_t(instance_blocks(p.n)) * 64 * sizeof(float);
}

// Retires the requests that snapshot `sn` shows finished (and that still hold
// the slot they held when it was taken).
static void server_retire(sr_server& S, ServerSnap& sn) {
  if (!sn.pending) return;
  sn.pending = false;
  const auto now = std::chrono::steady_clock::now();
  std::vector<std::unique_ptr<ServerRequest>> still;
  std::lock_guard<std::mutex> lk(S.mu);
  for (auto& rp : S.active) {
    ServerRequest& r = *rp;
    const InstState& st = sn.states[r.slot];
    if (sn.ticket[r.slot] != r.ticket || st.phase != 2) {
      still.push_back(std::move(rp));
      continue;
    }
    ServerDone d;
    d.code = state_to_result(st, S.gicp, r.src, r.map->tgt, &d.result, &d.msg);
    d.latency_s = std::chrono::duration<double>(now - r.submitted).count();
    S.done[r.ticket] = std::move(d);
    S.finished.push_back(r.ticket);
    S.bytes_active -= r.reserve;
    S.free_slots.push_back(r.slot);
  }
  S.active = std::move(still);
}

// Prepares the admitted sources. With a voxel leaf and k-NN covariances they
// are uploaded and prepared together as segments of flat buffers on one worker
// stream (one launch set for the group); otherwise each on the worker pool.
static void server_prepare(sr_server& S, std::vector<std::unique_ptr<ServerRequest>>& admit) {
  std::vector<Prepared> prepared(admit.size());
  std::shared_ptr<SegHolder> hold;
  Receipt rc;
  const bool seg = S.gicp.voxel_leaf > 0.0f && !(S.gicp.covariance_radius > 0.0f);
  if (seg) {
    Workers::get().run(1, [&](size_t, hipStream_t ws) {
      // device copies of the host columns, one flat buffer per axis
      std::vector<uint32_t> off(admit.size() + 1, 0);
      for (size_t i = 0; i < admit.size(); ++i) off[i + 1] = off[i] + uint32_t(admit[i]->x.size());
      const size_t total = off.back();
      DevBuf<float> dx(std::max<size_t>(total, 1)), dy(std::max<size_t>(total, 1)),
          dz(std::max<size_t>(total, 1));
      std::vector<sr_xyz_view> views(admit.size());
      std::vector<const sr_xyz_view*> vp(admit.size());
      std::vector<bool> is_target(admit.size(), false);
      std::vector<Prepared*> outs(admit.size());
      for (size_t i = 0; i < admit.size(); ++i) {
        ServerRequest& r = *admit[i];
        const size_t n = r.x.size(), o = off[i];
        SR_HIP(hipMemcpyAsync(dx.p + o, r.x.data(), 4 * n, hipMemcpyHostToDevice, ws));
        SR_HIP(hipMemcpyAsync(dy.p + o, r.y.data(), 4 * n, hipMemcpyHostToDevice, ws));
        SR_HIP(hipMemcpyAsync(dz.p + o, r.z.data(), 4 * n, hipMemcpyHostToDevice, ws));
        views[i] = sr_xyz_view{dx.p + o, dy.p + o, dz.p + o, n, SR_MEM_DEVICE};
        vp[i] = &views[i];
        outs[i] = &prepared[i];
      }
      hold = prepare_segmented(vp, is_target, S.gicp, outs, rc, ws);
      if (!hold) fail(SR_ERR_INVALID_ARGUMENT, "server preparation did not apply");
      SR_HIP(hipStreamSynchronize(ws));  // before the staging columns return to the cache
    });
  } else {
    std::vector<sr_xyz_view> views(admit.size());
    std::vector<const sr_xyz_view*> vp(admit.size());
    std::vector<sr_gicp_config> cfgs(admit.size(), S.gicp);
    for (size_t i = 0; i < admit.size(); ++i) {
      ServerRequest& r = *admit[i];
      views[i] = sr_xyz_view{r.x.data(), r.y.data(), r.z.data(), r.x.size(), SR_MEM_HOST};
      vp[i] = &views[i];
    }
    prepare_many(vp, cfgs, prepared, false, rc);
  }
  for (size_t i = 0; i < admit.size(); ++i) {
    admit[i]->src = std::move(prepared[i]);
    admit[i]->hold = hold;
  }
}

static void server_step(sr_server& S) {
  hipStream_t s = thread_stream();
  ServerSnap& sn = S.snaps[S.steps % kServerDepth];
  // The step kServerDepth back used this snapshot: wait for its copy, retire.
  if (sn.pending) {
    SR_HIP(hipEventSynchronize(sn.done));
    server_retire(S, sn);
  }
  // Newer snapshots that have landed retire their finished requests too.
  for (int k = 1; k < kServerDepth; ++k) {
    ServerSnap& o = S.snaps[(S.steps + k) % kServerDepth];
    if (o.pending && hipEventQuery(o.done) == hipSuccess) server_retire(S, o);
  }
  // Admission: FIFO while slots and budget allow.
  std::vector<std::unique_ptr<ServerRequest>> admit;
  {
    std::lock_guard<std::mutex> lk(S.mu);
    uint64_t bytes = S.bytes_active;
    while (!S.queue.empty() && S.active.size() + admit.size() < S.cfg.max_active &&
           bytes + S.queue.front()->reserve <= S.cfg.hbm_budget_bytes) {
      bytes += S.queue.front()->reserve;
      admit.push_back(std::move(S.queue.front()));
      S.queue.pop_front();
    }
    S.bytes_active = bytes;
  }
  // One pass over the active requests (asynchronous on this thread's stream).
  const uint32_t count = uint32_t(S.active.size());
  if (count) {
    std::vector<IterInstance> hi(count);
    std::vector<uint32_t> hbi;
    uint32_t begin = 0;
    uint64_t points = 0;
    for (uint32_t i = 0; i < count; ++i) {
      ServerRequest& r = *S.active[i];
      hi[i] = describe_instance(r.src, r.map->tgt, S.gicp, S.states.p + r.slot, r.partials.p,
                                r.match.p, r.budget.p, r.seen.p, begin);
      for (uint32_t b = 0; b < hi[i].nblocks; ++b) hbi.push_back(i);
      begin += hi[i].nblocks;
      points += r.src.n;
    }
    S.insts.resize(count);
    S.block_inst.resize(begin);
    upload(S.insts.p, hi.data(), sizeof(IterInstance) * count, s);
    upload(S.block_inst.p, hbi.data(), sizeof(uint32_t) * begin, s);
    const PassScratch ps = grow_scratch(S.pass_queue, S.pass_ctl, S.pass_done, points, begin, s);
    gicp_pass(S.insts.p, S.block_inst.p, begin, count, ps, hard_stage_blocks(points), s);
  }
  // Meanwhile: prepare the admitted sources on a worker stream.
  if (!admit.empty()) {
    try {
      server_prepare(S, admit);
    } catch (const Error& e) {
      // The whole admission group fails alike; report it on each ticket.
      std::lock_guard<std::mutex> lk(S.mu);
      for (auto& r : admit) {
        ServerDone d;
        d.code = e.code;
        d.msg = e.msg;
        S.done[r->ticket] = d;
        S.finished.push_back(r->ticket);
        S.bytes_active -= r->reserve;
      }
      admit.clear();
    }
    std::vector<double> init(7 * admit.size());
    for (size_t i = 0; i < admit.size(); ++i) {
      ServerRequest& r = *admit[i];
      // the budget charge becomes the measured footprint; later estimates
      // follow the largest bytes per raw point seen
      const uint64_t raw = std::max<uint64_t>(r.x.size(), 1);
      const uint64_t have = resident_bytes(r.src);
      S.bytes_per_point = std::max(S.bytes_per_point, double(have) / double(raw));
      {
        std::lock_guard<std::mutex> lk(S.mu);
        S.bytes_active = S.bytes_active - r.reserve + have;
      }
      r.reserve = have;
      std::vector<float>().swap(r.x);  // the host copy is no longer needed
      std::vector<float>().swap(r.y);
      std::vector<float>().swap(r.z);
      r.slot = S.free_slots.back();
      S.free_slots.pop_back();
      r.partials.resize(size_t(instance_blocks(r.src.n)) * 64);
      r.match.resize(std::max<uint32_t>(r.src.n, 1));
      fill_u32(r.match.p, kNone, r.src.n, s);
      r.budget.resize(std::max<uint32_t>(r.src.n, 1));
      fill_u32(reinterpret_cast<uint32_t*>(r.budget.p), 0u, r.src.n, s);
      r.seen.resize(std::max<uint32_t>(r.src.n, 1));
      fill_u32(reinterpret_cast<uint32_t*>(r.seen.p), 0xFFFFFFFFu, 4 * uint64_t(r.src.n), s);  // NaN
      for (int a = 0; a < 4; ++a) init[7 * i + a] = double(S.gicp.init_q[a]);
      for (int a = 0; a < 3; ++a) init[7 * i + 4 + a] = double(S.gicp.init_t[a]);
    }
    if (!admit.empty()) {
      S.init.resize(init.size());
      upload(S.init.p, init.data(), sizeof(double) * init.size(), s);
      for (size_t i = 0; i < admit.size(); ++i)
        init_states(S.states.p + admit[i]->slot, S.init.p + 7 * i, 1, s);
    }
  }
  for (auto& r : admit) S.active.push_back(std::move(r));
  // Snapshot of every slot after this step's work, copied asynchronously.
  if (!S.active.empty()) {
    std::fill(sn.ticket.begin(), sn.ticket.end(), 0ull);
    for (auto& r : S.active) sn.ticket[r->slot] = r->ticket;
    SR_HIP(hipMemcpyAsync(sn.states, S.states.p, sizeof(InstState) * S.states.n,
                          hipMemcpyDeviceToHost, s));
    SR_HIP(hipEventRecord(sn.done, s));
    sn.pending = true;
  }
  ++S.steps;
}

}  // namespace sr

// ---- scan-to-map odometry ------------------------------------------------

namespace sr {

// FP64 rigid pose (unit quaternion x y z w, translation) on the host.
struct Pose64 {
  double q[4] = {0.0, 0.0, 0.0, 1.0};
  double t[3] = {0.0, 0.0, 0.0};
};

static void quat_mul64(const double* a, const double* b, double* o) {
  o[0] = a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1];
  o[1] = a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0];
  o[2] = a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3];
  o[3] = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
}

static void quat_mat64(const double* q, double* r) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  r[0] = 1 - 2 * (y * y + z * z); r[1] = 2 * (x * y - z * w);     r[2] = 2 * (x * z + y * w);
  r[3] = 2 * (x * y + z * w);     r[4] = 1 - 2 * (x * x + z * z); r[5] = 2 * (y * z - x * w);
  r[6] = 2 * (x * z - y * w);     r[7] = 2 * (y * z + x * w);     r[8] = 1 - 2 * (x * x + y * y);
}

// a o b (apply b, then a), as Isometry3::compose (crates/spatialrust-math/src/transform.rs:104-108).
static Pose64 compose64(const Pose64& a, const Pose64& b) {
  Pose64 o;
  quat_mul64(a.q, b.q, o.q);
  const double n = std::sqrt(o.q[0] * o.q[0] + o.q[1] * o.q[1] + o.q[2] * o.q[2] + o.q[3] * o.q[3]);
  for (double& v : o.q) v /= n;
  double r[9];
  quat_mat64(a.q, r);
  for (int i = 0; i < 3; ++i)
    o.t[i] = r[3 * i] * b.t[0] + r[3 * i + 1] * b.t[1] + r[3 * i + 2] * b.t[2] + a.t[i];
  return o;
}

static Pose64 inverse64(const Pose64& a) {
  Pose64 o;
  o.q[0] = -a.q[0];
  o.q[1] = -a.q[1];
  o.q[2] = -a.q[2];
  o.q[3] = a.q[3];
  double r[9];
  quat_mat64(o.q, r);
  for (int i = 0; i < 3; ++i)
    o.t[i] = -(r[3 * i] * a.t[0] + r[3 * i + 1] * a.t[1] + r[3 * i + 2] * a.t[2]);
  return o;
}

}  // namespace sr

// Scan-to-map odometry: each scan is voxel-filtered, registered against the
// rolling map with a constant-velocity initial guess, and merged into the map.
struct sr_odometry {
  sr_odometry_config cfg{};
  sr_gicp_config gicp{};
  sr::VoxelMap map;
  sr::Prepared map_prep;  // grid, empty field and normals of map.list
  sr::Pose64 pose, prev;  // map_T_scan of the last two scans
  uint64_t scans = 0;
  // Grid geometry of the map (bounds as ordered words, cell edge), kept from
  // one full rebuild to the next so unchanged neighbourhoods keep their normals.
  bool geo_valid = false;
  uint32_t geo_hb[6] = {};
  float geo_h = 0.0f;
  uint32_t geo_points = 0;  // map size at the last full rebuild
  uint64_t geo_drops = 0;   // map.drops at the last full rebuild
  sr::DevBuf<float4> lnrm;  // normals of the map points, list order
  sr::DevBuf<float> ltau;   // k-th neighbour squared distance of each, list order
  sr::DevBuf<uint32_t> refresh, refresh_n;
  uint64_t refreshed = 0, rebuilds = 0;  // normals computed by the last push; full rebuilds
};

template <typename F>
static int32_t guarded(F&& f) {
  try {
    t_last_error.clear();
    f();
    return SR_OK;
  } catch (const Error& e) {
    t_last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    t_last_error = "host allocation failed";
    return SR_ERR_INVALID_ARGUMENT;
  } catch (const std::exception& e) {
    t_last_error = e.what();
    return SR_ERR_INVALID_ARGUMENT;
  }
}

extern "C" {

const char* sr_last_error(void) { return t_last_error.c_str(); }

void sr_profile_enable(int32_t on) { profiler().on.store(on == 2 ? 2 : (on ? 1 : 0)); }

int32_t sr_profile_read_stats(const char* name, uint64_t* out10) {
  return guarded([&] {
    Profiler& P = profiler();
    SR_HIP(hipDeviceSynchronize());
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.stats.find(name);
    SearchStats h{};
    if (it != P.stats.end()) {
      SR_HIP(hipMemcpy(&h, it->second, sizeof(h), hipMemcpyDeviceToHost));
      SR_HIP(hipMemset(it->second, 0, sizeof(h)));
    }
    out10[0] = h.queries;
    out10[1] = h.rings;
    out10[2] = h.cells;
    out10[3] = h.points;
    for (int o = 0; o < 4; ++o) out10[4 + o] = h.outcome[o];
    out10[8] = h.max_cells;
    out10[9] = h.max_points;
  });
}

void sr_profile_reset(void) {
  Profiler& P = profiler();
  (void)hipDeviceSynchronize();
  std::lock_guard<std::mutex> lk(P.mu);
  for (auto& r : P.recs) {
    P.spare.push_back(r.a);
    P.spare.push_back(r.b);
  }
  P.recs.clear();
}

int32_t sr_profile_read(const char* name, double* total_ms, uint64_t* launches) {
  return guarded([&] {
    Profiler& P = profiler();
    SR_HIP(hipDeviceSynchronize());
    std::lock_guard<std::mutex> lk(P.mu);
    double tot = 0.0;
    uint64_t cnt = 0;
    for (auto& r : P.recs) {
      if (r.name != name) continue;
      float ms = 0.0f;
      SR_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      tot += ms;
      ++cnt;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = cnt;
  });
}

int64_t sr_profile_read_each(const char* name, double* out_ms, uint64_t cap) {
  Profiler& P = profiler();
  if (hipDeviceSynchronize() != hipSuccess) return -1;
  std::lock_guard<std::mutex> lk(P.mu);
  uint64_t cnt = 0;
  for (auto& r : P.recs) {
    if (r.name != name) continue;
    if (cnt < cap && out_ms) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) return -1;
      out_ms[cnt] = ms;
    }
    ++cnt;
  }
  return int64_t(cnt);
}

}  // extern "C"

// The 0/1 switches among the options.
static std::atomic<int64_t>* switch_option(const std::string& n) {
  Options& o = options();
  if (n == "match_reuse") return &o.match_reuse;
  if (n == "radius_tiles") return &o.radius_tiles;
  if (n == "autotune") return &o.autotune;
  if (n == "odometry_incremental") return &o.odometry_incremental;
  if (n == "pass_graphs") return &o.pass_graphs;
  return nullptr;
}

extern "C" {

int32_t sr_option_set(const char* name, int64_t value) {
  return guarded([&] {
    if (!name) fail(SR_ERR_INVALID_ARGUMENT, "option name must not be null");
    const std::string n(name);
    if (std::atomic<int64_t>* sw = switch_option(n)) {
      if (value != 0 && value != 1) fail(SR_ERR_INVALID_ARGUMENT, n + " takes 0 or 1");
      sw->store(value);
    } else if (n == "tune_force") {
      if (value < -1) fail(SR_ERR_INVALID_ARGUMENT, "tune_force takes -1 or a candidate index");
      options().tune_force.store(value);
    } else {
      fail(SR_ERR_INVALID_ARGUMENT, "unknown option " + n);
    }
  });
}

int32_t sr_option_get(const char* name, int64_t* value) {
  return guarded([&] {
    if (!name || !value) fail(SR_ERR_INVALID_ARGUMENT, "null option argument");
    const std::string n(name);
    if (std::atomic<int64_t>* sw = switch_option(n)) {
      *value = sw->load();
    } else if (n == "tune_force") {
      *value = options().tune_force.load();
    } else if (n == "tune_candidates_single") {
      *value = int64_t(tune_candidates(false).size());
    } else if (n == "tune_candidates_batch") {
      *value = int64_t(tune_candidates(true).size());
    } else if (n == "noalloc_scratch_bytes") {
      *value = int64_t(kScratchBytes);
    } else if (n == "noalloc_peak_bytes") {  // of the last allocation-free call on this device
      int dev = 0;
      SR_HIP(hipGetDevice(&dev));
      ScratchDevice& sd = scratch_device(dev);
      std::lock_guard<std::mutex> lk(sd.call);
      *value = int64_t(sd.last_peak);
    } else {
      fail(SR_ERR_INVALID_ARGUMENT, "unknown option " + n);
    }
  });
}

int32_t sr_tune_describe(char* out, uint64_t cap, uint64_t* len) {
  return guarded([&] {
    std::string text;
    {
      Tuner& T = tuner();
      std::lock_guard<std::mutex> lk(T.mu);
      for (const auto& kv : T.done) {
        const PassTuning& b = kv.second.best;
        text += kv.first.first + " class " + std::string(kv.first.second & 0x100u ? "batch" : "single") +
                " log2_points " + std::to_string(kv.first.second & 0xFFu) + ": group " +
                std::to_string(b.group) + " hard_cap " + std::to_string(b.hard_cap) + " split " +
                std::to_string(b.split_blocks) + " ms";
        for (float m : kv.second.ms) text += " " + std::to_string(m);
        text += "\n";
      }
    }
    if (len) *len = text.size();
    if (out && cap) {
      const size_t n = std::min<size_t>(cap - 1, text.size());
      std::memcpy(out, text.data(), n);
      out[n] = '\0';
    }
  });
}

int32_t sr_device_available(void) { return device_state().status == 0 ? 1 : 0; }

int32_t sr_version(void) { return 100; }

void sr_gicp_config_default(sr_gicp_config* out) {
  if (!out) return;
  std::memset(out, 0, sizeof(*out));
  out->max_iterations = 50;
  out->max_correspondence_distance = 1.0f;
  out->k_neighbors = 20;
  out->epsilon = 1e-3;
  out->transformation_epsilon = 1e-8;
  out->fitness_epsilon = 1e-6;
  out->min_correspondences = 6;
  out->init_q[3] = 1.0f;
  out->covariance_radius = 0.0f;
  out->voxel_leaf = 0.0f;
}

int32_t sr_gicp_align(const sr_xyz_view* source, const sr_xyz_view* target,
                      const sr_gicp_config* config, sr_policy policy, sr_result* out,
                      sr_receipt* receipt) {
  return guarded([&] {
    if (receipt) {
      std::memset(receipt, 0, sizeof(*receipt));
      receipt->requested = policy;
    }
    sr_policy resolved = resolve(policy);
    validate_view(source);
    validate_view(target);
    validate_cfg(source, target, config);
    hipStream_t s = thread_stream();
    Receipt rc;
    Prepared tgt, src;
    // device views with a voxel leaf and k-NN covariances: both clouds as two
    // segments of one preparation (its host waits shared by the two clouds);
    // the results are those of preparing each alone
    std::vector<Prepared*> outs = {&tgt, &src};
    const std::shared_ptr<SegHolder> hold =
        prepare_segmented({target, source}, {true, false}, *config, outs, rc, s);
    if (!hold) {
      prepare(*target, *config, tgt, rc, s, true);
      prepare(*source, *config, src, rc, s, false);
    }
    Batch B;
    batch_setup({&src}, {&tgt}, {*config}, B, s);
    batch_run(B, s);
    rc.stages |= SR_STAGE_ITERATE;
    InstState st;
    download(&st, B.states.p, sizeof(st), s);
    rc.d2h += sizeof(st);
    rc.stages |= SR_STAGE_READBACK;
    if (receipt) {
      receipt->resolved = resolved;
      receipt->h2d_bytes = rc.h2d;
      receipt->d2h_bytes = rc.d2h;
      receipt->d2d_bytes = rc.d2d;
      receipt->stage_mask = rc.stages;
    }
    std::string msg;
    int32_t code = state_to_result(st, *config, src, tgt, out, &msg);
    if (code) fail(code, msg);
  });
}

int32_t sr_gicp_align_noalloc(const sr_xyz_view* source, const sr_xyz_view* target,
                              const sr_gicp_config* config, sr_result* out) {
  return guarded([&] {
    resolve(sr_policy{SR_GPU, SR_DEV_HIP});
    validate_view(source);
    validate_view(target);
    validate_cfg(source, target, config);
    if (source->n > kNoAllocMaxPoints || target->n > kNoAllocMaxPoints)
      fail(SR_ERR_INVALID_ARGUMENT,
           "sr_gicp_align_noalloc takes clouds of at most " + std::to_string(kNoAllocMaxPoints) +
               " points (got " + std::to_string(source->n) + " and " + std::to_string(target->n) + ")");
    hipStream_t s = thread_stream();
    int dev = 0;
    SR_HIP(hipGetDevice(&dev));
    ScratchDevice& sd = scratch_device(dev);
    std::lock_guard<std::mutex> lk(sd.call);
    ArenaScope scope(sd, s);  // declared before every buffer of the call
    Receipt rc;
    Prepared tgt, src;
    prepare(*target, *config, tgt, rc, s, true);
    prepare(*source, *config, src, rc, s, false);
    Batch B;
    batch_setup({&src}, {&tgt}, {*config}, B, s);
    batch_run(B, s);
    InstState st;
    download(&st, B.states.p, sizeof(st), s);
    std::string msg;
    const int32_t code = state_to_result(st, *config, src, tgt, out, &msg);
    if (code) fail(code, msg);
  });
}

int32_t sr_gicp_normal_equations(const sr_xyz_view* source, const sr_xyz_view* target,
                                 const sr_gicp_config* config, const double* q, const double* t,
                                 int32_t* out_match, double* out_H, double* out_g,
                                 uint64_t* out_count, double* out_d2_sum) {
  return guarded([&] {
    resolve(sr_policy{SR_GPU, SR_DEV_HIP});
    validate_view(source);
    validate_view(target);
    validate_cfg(source, target, config);
    if (!q || !t) fail(SR_ERR_INVALID_ARGUMENT, "pose must not be null");
    sr_gicp_config cfg = *config;
    cfg.voxel_leaf = 0.0f;  // clouds as given, so indices refer to the caller's points
    hipStream_t s = thread_stream();
    Receipt rc;
    Prepared tgt, src;
    prepare(*target, cfg, tgt, rc, s, true);
    prepare(*source, cfg, src, rc, s, false);
    double pose[7] = {q[0], q[1], q[2], q[3], t[0], t[1], t[2]};
    Batch B;
    batch_setup({&src}, {&tgt}, {cfg}, B, s, pose);
    gicp_pass(B.insts.p, B.block_inst.p, B.total_blocks, 1, B.scratch, B.hard_blocks, s, false);
    std::vector<float> part(size_t(B.total_blocks) * 64);
    std::vector<uint32_t> match(src.n), sorder(src.n), torder(tgt.n);
    SR_HIP(hipMemcpyAsync(part.data(), B.partials.p, part.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    SR_HIP(hipMemcpyAsync(match.data(), B.match.p, match.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    SR_HIP(hipMemcpyAsync(sorder.data(), src.grid.order.p, sorder.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    download(torder.data(), tgt.grid.order.p, torder.size() * sizeof(uint32_t), s);
    // Block partials in block order, each term as its (hi, lo) pair, in FP64:
    // the same sums the finalize kernel forms.
    double tot[kTerms] = {};
    uint64_t cnt = 0;
    for (uint32_t b = 0; b < B.total_blocks; ++b) {
      const float* pb = part.data() + size_t(b) * 64;
      for (int k = 0; k < kTerms; ++k) tot[k] += double(pb[2 * k]) + double(pb[2 * k + 1]);
      uint64_t c;
      std::memcpy(&c, pb + 60, sizeof(c));
      cnt += c;
    }
    if (out_H)
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) out_H[6 * a + b] = tot[a <= b ? ut(a, b) : ut(b, a)];
    if (out_g)
      for (int a = 0; a < 6; ++a) out_g[a] = tot[21 + a];
    if (out_count) *out_count = cnt;
    if (out_d2_sum) *out_d2_sum = tot[27];
    if (out_match)
      for (uint32_t i = 0; i < src.n; ++i)
        out_match[sorder[i]] = match[i] == kNone ? -1 : int32_t(torder[match[i]]);
  });
}

int32_t sr_voxel_downsample(const sr_xyz_view* input, float leaf, const float* origin,
                            sr_policy policy, float* out_x, float* out_y, float* out_z,
                            int64_t* keys, uint32_t* counts, uint64_t* n_out, sr_receipt* receipt) {
  return guarded([&] {
    if (n_out) *n_out = 0;
    if (receipt) {
      std::memset(receipt, 0, sizeof(*receipt));
      receipt->requested = policy;
    }
    sr_policy resolved = resolve(policy);
    validate_view(input);
    if (input->n == 0) return;  // voxel.rs:257-259
    if (!(leaf > 0.0f)) fail(SR_ERR_INVALID_ARGUMENT, "leaf_size must be greater than zero");
    if (input->n > 0xFFFFFFF0ull) fail(SR_ERR_INVALID_ARGUMENT, "point clouds above 4G points are not supported");
    hipStream_t s = thread_stream();
    Receipt rc;
    Cloud raw;
    pack_cloud(input->x, input->y, input->z, uint32_t(input->n), input->mem == SR_MEM_DEVICE, raw, s,
               &rc.h2d);
    VoxelOut vox;
    voxel_downsample(raw, leaf, origin, keys != nullptr, vox, s);
    const uint32_t c = vox.cells;
    std::vector<float4> host(c);
    if (c) SR_HIP(hipMemcpyAsync(host.data(), vox.centroids.pts.p, sizeof(float4) * c, hipMemcpyDeviceToHost, s));
    if (keys && c)
      SR_HIP(hipMemcpyAsync(keys, vox.keys.p, sizeof(int64_t) * 3 * c, hipMemcpyDeviceToHost, s));
    if (counts && c)
      SR_HIP(hipMemcpyAsync(counts, vox.counts.p, sizeof(uint32_t) * c, hipMemcpyDeviceToHost, s));
    SR_HIP(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < c; ++i) {
      out_x[i] = host[i].x;
      out_y[i] = host[i].y;
      out_z[i] = host[i].z;
    }
    rc.d2h += uint64_t(12) * c;
    if (n_out) *n_out = c;
    if (receipt) {
      receipt->resolved = resolved;
      receipt->h2d_bytes = rc.h2d;
      receipt->d2h_bytes = rc.d2h;
      receipt->stage_mask = SR_STAGE_VOXEL | SR_STAGE_READBACK;
    }
  });
}

int32_t sr_plane_covariances(const sr_xyz_view* cloud, uint32_t k_neighbors, float radius,
                             double epsilon, float* out_cov6, float* out_normals) {
  return guarded([&] {
    ensure_device();
    validate_view(cloud);
    if (cloud->n == 0) return;
    if (radius <= 0.0f && k_neighbors < 3)
      fail(SR_ERR_INVALID_ARGUMENT, "GICP needs k_neighbors >= 3 for covariance estimation");
    if (radius > 0.0f && !(radius == radius)) fail(SR_ERR_INVALID_ARGUMENT, "grid radius must be positive");
    hipStream_t s = thread_stream();
    Cloud c;
    pack_cloud(cloud->x, cloud->y, cloud->z, uint32_t(cloud->n), cloud->mem == SR_MEM_DEVICE, c, s, nullptr);
    uint32_t hb[6];
    host_bounds(c, hb, s);
    const uint32_t n = c.n;
    DevBuf<float> cov6(size_t(6) * n), nrm3(size_t(3) * n);
    if (radius > 0.0f) {
      Grid rg;
      build_grid(c, hb, radius, rg, s);
      radius_normals(rg, radius, float(epsilon), nullptr, cov6.p, nrm3.p, s);
    } else {
      Grid g;
      build_grid(c, hb, choose_cell_edge(c, hb, knn_occupancy(k_neighbors)), g, s);
      DevBuf<float4> rec(size_t(kRec) * n);
      knn_normals(g, k_neighbors, epsilon, rec.p, cov6.p, nrm3.p, s);
    }
    if (out_cov6) SR_HIP(hipMemcpyAsync(out_cov6, cov6.p, sizeof(float) * 6 * n, hipMemcpyDeviceToHost, s));
    if (out_normals) SR_HIP(hipMemcpyAsync(out_normals, nrm3.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
    SR_HIP(hipStreamSynchronize(s));
  });
}

int32_t sr_grid_normals(const sr_xyz_view* cloud, float radius, const float* viewpoint,
                        float* out_normals, float* out_curvature) {
  return guarded([&] {
    ensure_device();
    validate_view(cloud);
    if (cloud->n == 0) return;  // normal_gpu.rs:35-37
    if (!(radius > 0.0f)) fail(SR_ERR_INVALID_ARGUMENT, "grid radius must be positive");
    if (cloud->n > 0xFFFFFFF0ull) fail(SR_ERR_INVALID_ARGUMENT, "point clouds above 4G points are not supported");
    hipStream_t s = thread_stream();
    Cloud c;
    pack_cloud(cloud->x, cloud->y, cloud->z, uint32_t(cloud->n), cloud->mem == SR_MEM_DEVICE, c, s, nullptr);
    uint32_t hb[6];
    host_bounds(c, hb, s);
    Grid rg;
    build_grid(c, hb, radius, rg, s);
    const uint32_t n = c.n;
    DevBuf<float4> out(n);
    grid_normals(rg, radius, viewpoint, out.p, s);
    std::vector<float4> host(n);
    download(host.data(), out.p, sizeof(float4) * n, s);
    for (uint32_t i = 0; i < n; ++i) {
      if (out_normals) {
        out_normals[3 * size_t(i)] = host[i].x;
        out_normals[3 * size_t(i) + 1] = host[i].y;
        out_normals[3 * size_t(i) + 2] = host[i].z;
      }
      if (out_curvature) out_curvature[i] = host[i].w;
    }
  });
}

int32_t sr_knn(const sr_xyz_view* points, const sr_xyz_view* queries, uint32_t k,
               int32_t* out_index, float* out_d2) {
  return guarded([&] {
    ensure_device();
    validate_view(points);
    validate_view(queries);
    if (queries->n == 0 || k == 0) return;
    hipStream_t s = thread_stream();
    if (points->n == 0) {
      for (uint64_t i = 0; i < queries->n * k; ++i) {
        out_index[i] = -1;
        out_d2[i] = -1.0f;
      }
      return;
    }
    Cloud pc, qc;
    pack_cloud(points->x, points->y, points->z, uint32_t(points->n), points->mem == SR_MEM_DEVICE, pc, s, nullptr);
    pack_cloud(queries->x, queries->y, queries->z, uint32_t(queries->n), queries->mem == SR_MEM_DEVICE, qc, s,
               nullptr);
    uint32_t hb[6];
    host_bounds(pc, hb, s);
    Grid g;
    build_grid(pc, hb, choose_cell_edge(pc, hb, knn_occupancy(k)), g, s);
    DevBuf<int32_t> idx(size_t(qc.n) * k);
    DevBuf<float> d2(size_t(qc.n) * k);
    knn_query(g, qc.pts.p, qc.n, k, idx.p, d2.p, s);
    SR_HIP(hipMemcpyAsync(out_index, idx.p, sizeof(int32_t) * qc.n * k, hipMemcpyDeviceToHost, s));
    download(out_d2, d2.p, sizeof(float) * qc.n * k, s);
  });
}

int32_t sr_map_create(const sr_xyz_view* target, const sr_gicp_config* config, sr_map** out) {
  return guarded([&] {
    if (!out) fail(SR_ERR_INVALID_ARGUMENT, "out must not be null");
    *out = nullptr;
    ensure_device();
    validate_view(target);
    validate_cfg(target, target, config);
    hipStream_t s = thread_stream();
    auto m = std::make_unique<sr_map>();
    m->cfg = *config;
    Receipt rc;
    prepare(*target, *config, m->tgt, rc, s, true);
    SR_HIP(hipStreamSynchronize(s));
    *out = m.release();
  });
}

void sr_map_destroy(sr_map* map) {
  if (!map) return;
  (void)hipDeviceSynchronize();
  delete map;
}

uint64_t sr_map_points(const sr_map* map) { return map ? map->tgt.n : 0; }

uint64_t sr_map_device_bytes(const sr_map* map) {
  if (!map) return 0;
  const Prepared& p = map->tgt;
  return p.grid.cell_start.bytes() + p.grid.pts.bytes() + p.grid.order.bytes() + p.rec.bytes();
}

int32_t sr_map_align(const sr_map* map, const sr_xyz_view* source, const sr_gicp_config* config,
                     sr_result* out) {
  return guarded([&] {
    if (!map) fail(SR_ERR_INVALID_ARGUMENT, "map must not be null");
    ensure_device();
    validate_view(source);
    const sr_xyz_view dummy{nullptr, nullptr, nullptr, map->tgt.n, 0};
    validate_cfg(source, &dummy, config);
    check_map_config(map->cfg, *config);
    hipStream_t s = thread_stream();
    Receipt rc;
    Prepared src;
    prepare(*source, *config, src, rc, s, false);
    Batch B;
    batch_setup({&src}, {&map->tgt}, {*config}, B, s);
    batch_run(B, s);
    InstState st;
    download(&st, B.states.p, sizeof(st), s);
    std::string msg;
    int32_t code = state_to_result(st, *config, src, map->tgt, out, &msg);
    if (code) fail(code, msg);
  });
}

int32_t sr_batch_align(const sr_map* const* maps, const sr_xyz_view* sources, uint32_t count,
                       const sr_gicp_config* config, sr_result* out) {
  return guarded([&] {
    if (count == 0) return;
    if (!maps || !sources || !out) fail(SR_ERR_INVALID_ARGUMENT, "null batch argument");
    ensure_device();
    hipStream_t s = thread_stream();
    std::vector<Prepared> srcs(count);
    std::vector<const Prepared*> sp(count), tp(count);
    std::vector<const sr_xyz_view*> views(count);
    std::vector<sr_gicp_config> cfgs(count, *config);
    Receipt rc;
    for (uint32_t i = 0; i < count; ++i) {
      if (!maps[i]) fail(SR_ERR_INVALID_ARGUMENT, "map must not be null");
      validate_view(&sources[i]);
      const sr_xyz_view dummy{nullptr, nullptr, nullptr, maps[i]->tgt.n, 0};
      validate_cfg(&sources[i], &dummy, config);
      check_map_config(maps[i]->cfg, *config);
      views[i] = &sources[i];
      sp[i] = &srcs[i];
      tp[i] = &maps[i]->tgt;
    }
    prepare_many(views, cfgs, srcs, false, rc);
    Batch B;
    batch_setup(sp, tp, cfgs, B, s);
    batch_run(B, s);
    std::vector<InstState> st(count);
    download(st.data(), B.states.p, sizeof(InstState) * count, s);
    int32_t first_code = 0;
    std::string first_msg;
    for (uint32_t i = 0; i < count; ++i) {
      std::string msg;
      int32_t code = state_to_result(st[i], *config, srcs[i], maps[i]->tgt, &out[i], &msg);
      if (code && !first_code) {
        first_code = code;
        first_msg = "registration " + std::to_string(i) + ": " + msg;
      }
    }
    if (first_code) fail(first_code, first_msg);
  });
}

int32_t sr_batch_align_pairs(const sr_xyz_view* sources, const sr_xyz_view* targets,
                             uint32_t count, const sr_gicp_config* config, sr_result* out) {
  return guarded([&] {
    if (count == 0) return;
    if (!sources || !targets || !out || !config) fail(SR_ERR_INVALID_ARGUMENT, "null batch argument");
    ensure_device();
    for (uint32_t i = 0; i < count; ++i) {
      validate_view(&sources[i]);
      validate_view(&targets[i]);
      validate_cfg(&sources[i], &targets[i], config);
    }
    hipStream_t s = thread_stream();
    std::vector<Prepared> srcs(count), tgts(count);
    std::vector<const sr_xyz_view*> sv(count), tv(count);
    std::vector<const Prepared*> sp(count), tp(count);
    std::vector<sr_gicp_config> cfgs(count, *config);
    for (uint32_t i = 0; i < count; ++i) {
      sv[i] = &sources[i];
      tv[i] = &targets[i];
      sp[i] = &srcs[i];
      tp[i] = &tgts[i];
    }
    Receipt rc;
    // the prepared clouds may view into `hold`'s buffers: keep it to the end
    const std::shared_ptr<SegHolder> hold = prepare_pairs(tv, sv, cfgs, tgts, srcs, rc);
    Batch B;
    batch_setup(sp, tp, cfgs, B, s);
    batch_run(B, s);
    std::vector<InstState> st(count);
    download(st.data(), B.states.p, sizeof(InstState) * count, s);
    int32_t first_code = 0;
    std::string first_msg;
    for (uint32_t i = 0; i < count; ++i) {
      std::string msg;
      int32_t code = state_to_result(st[i], *config, srcs[i], tgts[i], &out[i], &msg);
      if (code && !first_code) {
        first_code = code;
        first_msg = "registration " + std::to_string(i) + ": " + msg;
      }
    }
    if (first_code) fail(first_code, first_msg);
  });
}

int32_t sr_server_create(const sr_server_config* config, const sr_gicp_config* gicp,
                         sr_server** out) {
  return guarded([&] {
    if (!out) fail(SR_ERR_INVALID_ARGUMENT, "out must not be null");
    *out = nullptr;
    if (!config || !gicp) fail(SR_ERR_INVALID_ARGUMENT, "server config must not be null");
    if (config->max_active == 0) fail(SR_ERR_INVALID_ARGUMENT, "max_active must be at least 1");
    check_k(gicp->k_neighbors, gicp->covariance_radius);
    ensure_device();
    auto sv = std::make_unique<sr_server>();
    sv->cfg = *config;
    sv->gicp = request_config(*config, *gicp);
    sv->states.resize(config->max_active);
    for (uint32_t i = config->max_active; i > 0; --i) sv->free_slots.push_back(i - 1);
    for (auto& sn : sv->snaps) {
      SR_HIP(hipHostMalloc(reinterpret_cast<void**>(&sn.states),
                           sizeof(InstState) * config->max_active, hipHostMallocDefault));
      SR_HIP(hipEventCreateWithFlags(&sn.done, hipEventDisableTiming));
      sn.ticket.assign(config->max_active, 0ull);
    }
    *out = sv.release();
  });
}

void sr_server_destroy(sr_server* server) {
  if (!server) return;
  (void)hipDeviceSynchronize();
  delete server;
}

int32_t sr_server_submit(sr_server* server, const sr_map* map, const sr_xyz_view* source,
                         uint64_t* ticket) {
  return guarded([&] {
    if (!server || !map || !ticket) fail(SR_ERR_INVALID_ARGUMENT, "null server argument");
    validate_view(source);
    const sr_xyz_view dummy{nullptr, nullptr, nullptr, map->tgt.n, 0};
    validate_cfg(source, &dummy, &server->gicp);
    check_map_config(map->cfg, server->gicp);
    if (source->mem != SR_MEM_HOST)
      fail(SR_ERR_INVALID_ARGUMENT, "server requests take host xyz columns");
    // admission charge until the request is prepared and measured: its
    // preparation's transient buffers (kPrepBytesPerPoint) on top of the largest
    // resident bytes per raw point measured so far (a 64-byte floor)
    const uint64_t need =
        uint64_t((server->bytes_per_point + kPrepBytesPerPoint) * double(source->n)) + (1u << 20);
    if (need > server->cfg.hbm_budget_bytes)
      fail(SR_ERR_INVALID_ARGUMENT, "request needs " + std::to_string(need) +
                                        " device bytes, above the server's HBM budget");
    auto r = std::make_unique<ServerRequest>();
    r->map = map;
    r->x.assign(source->x, source->x + source->n);
    r->y.assign(source->y, source->y + source->n);
    r->z.assign(source->z, source->z + source->n);
    r->reserve = need;
    r->submitted = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(server->mu);
    r->ticket = server->next_ticket++;
    *ticket = r->ticket;
    server->queue.push_back(std::move(r));
  });
}

int32_t sr_server_step(sr_server* server, uint32_t* pending) {
  return guarded([&] {
    if (!server) fail(SR_ERR_INVALID_ARGUMENT, "server must not be null");
    server_step(*server);
    if (pending) {
      std::lock_guard<std::mutex> lk(server->mu);
      *pending = uint32_t(server->queue.size() + server->active.size());
    }
  });
}

int32_t sr_server_poll(sr_server* server, uint64_t ticket, sr_result* out, uint8_t* ready) {
  return guarded([&] {
    if (!server || !ready) fail(SR_ERR_INVALID_ARGUMENT, "null server argument");
    *ready = 0;
    std::lock_guard<std::mutex> lk(server->mu);
    auto it = server->done.find(ticket);
    if (it == server->done.end()) return;
    *ready = 1;
    if (out) *out = it->second.result;
    const int32_t code = it->second.code;
    const std::string msg = it->second.msg;
    server->done.erase(it);
    if (code) fail(code, msg);
  });
}

int32_t sr_server_latency(sr_server* server, uint64_t ticket, double* seconds) {
  return guarded([&] {
    if (!server || !seconds) fail(SR_ERR_INVALID_ARGUMENT, "null server argument");
    std::lock_guard<std::mutex> lk(server->mu);
    auto it = server->done.find(ticket);
    if (it == server->done.end()) fail(SR_ERR_INVALID_ARGUMENT, "ticket has not finished");
    *seconds = it->second.latency_s;
  });
}

int32_t sr_server_finished(sr_server* server, uint64_t* tickets, uint64_t cap, uint64_t* n) {
  return guarded([&] {
    if (!server || !n) fail(SR_ERR_INVALID_ARGUMENT, "null server argument");
    *n = 0;
    if (cap && !tickets) fail(SR_ERR_INVALID_ARGUMENT, "tickets must not be null");
    std::lock_guard<std::mutex> lk(server->mu);
    while (*n < cap && !server->finished.empty()) {
      tickets[(*n)++] = server->finished.front();
      server->finished.pop_front();
    }
  });
}

uint64_t sr_server_device_bytes(const sr_server* server) {
  return server ? server->bytes_active : 0;
}

}  // extern "C"

// Grid, empty field and normals of the odometry map after points [n_old,
// list.n) were appended (n_old = 0 after a drop renumbered the list). A full
// rebuild chooses the grid geometry: the list's bounds padded by an eighth of
// their span plus 16 leaves per side, and the cell edge a fresh grid of the
// list would get. Until the next full rebuild (a drop, a new point outside the
// box, or the map doubling) every scan re-sorts the grid with that geometry and
// recomputes the k-NN normals of map_refresh_set's points only; the others keep
// theirs, which equal what a full computation on that geometry gives.
static void odometry_update_map(sr_odometry* o, uint32_t n_old, Receipt& rc, hipStream_t s) {
  const Cloud& list = o->map.list;
  const uint32_t n = list.n;
  if (o->gicp.covariance_radius > 0.0f) {  // radius covariances: the whole map each scan
    Prepared fresh;
    prepare_cloud(list, o->gicp, fresh, rc, s, true, o->cfg.voxel_leaf);
    SR_HIP(hipStreamSynchronize(s));
    o->map_prep = std::move(fresh);
    o->refreshed = n;
    ++o->rebuilds;
    return;
  }
  bool full = !o->geo_valid || o->map.drops != o->geo_drops || n_old == 0 ||
              uint64_t(n) >= 2ull * o->geo_points;
  if (!full && n > n_old) {
    DevBuf<uint32_t> b(8);
    uint32_t nb[6];
    cloud_bounds(list.pts.p + n_old, n - n_old, b.p, s);
    download(nb, b.p, sizeof(nb), s);
    for (int a = 0; a < 3; ++a) full = full || nb[a] < o->geo_hb[a] || nb[3 + a] > o->geo_hb[3 + a];
  }
  const float leaf = o->cfg.voxel_leaf;
  if (full) {
    uint32_t hb[6];
    host_bounds(list, hb, s);
    float h = choose_cell_edge(list, hb, knn_occupancy(o->gicp.k_neighbors), leaf);
    double span_cells[3];
    for (int a = 0; a < 3; ++a) {
      const float lo = ord2f(hb[a]), hi = ord2f(hb[3 + a]);
      const float pad = 0.125f * (hi - lo) + 16.0f * leaf;
      o->geo_hb[a] = f2ord(lo - pad);
      o->geo_hb[3 + a] = f2ord(hi + pad);
      span_cells[a] = double(ord2f(o->geo_hb[3 + a]) - ord2f(o->geo_hb[a]));
    }
    // the padded box stays within a quarter of the 64M-cell grid cap
    auto cells = [&](float e) {
      double c = 1.0;
      for (int a = 0; a < 3; ++a) c *= std::floor(span_cells[a] / double(e)) + 1.0;
      return c;
    };
    while (cells(h) > 16.0e6) h *= 1.125f;
    o->geo_h = h;
    o->geo_valid = true;
    o->geo_points = n;
    o->geo_drops = o->map.drops;
    ++o->rebuilds;
    // list-order buffers for every map size up to the next full rebuild (< 2n)
    const size_t cap = std::max<size_t>(std::min<uint64_t>(2ull * n, o->cfg.max_map_points), 1);
    if (o->lnrm.n < cap) {
      o->lnrm.resize(cap);
      o->ltau.resize(cap);
      o->refresh.resize(cap);
    }
    o->refresh_n.resize(1);
  }
  Prepared fresh;
  fresh.n = n;
  build_grid(list, o->geo_hb, o->geo_h, fresh.grid, s);
  build_empty_radius(fresh.grid, s);
  rc.stages |= SR_STAGE_GRID;
  fresh.rec.resize(kRec * std::max<uint32_t>(n, 1));
  KnnCloud kc = knn_cloud(fresh.grid, fresh.rec.p, nullptr, nullptr);
  kc.nrm_orig = o->lnrm.p;
  kc.tau_orig = o->ltau.p;
  const bool incremental = !full && options().odometry_incremental.load() != 0;
  if (incremental) {
    map_refresh_set(fresh.grid, list.pts.p, n_old, o->ltau.p, o->refresh.p, o->refresh_n.p, s);
    kc.subset = o->refresh.p;
    kc.subset_n = o->refresh_n.p;
  }
  knn_normals_many({kc}, o->gicp.k_neighbors, o->gicp.epsilon, s);
  if (incremental) gather_list_normals(fresh.grid, o->lnrm.p, fresh.rec.p, s);
  rc.stages |= SR_STAGE_COVARIANCE;
  uint32_t refreshed = n;
  if (incremental) download(&refreshed, o->refresh_n.p, sizeof(refreshed), s);
  else SR_HIP(hipStreamSynchronize(s));
  o->refreshed = refreshed;
  o->map_prep = std::move(fresh);
}

// One scan on its way through an odometry push.
struct OdoScan {
  VoxelOut vox;
  Prepared src;
  Pose64 pose;  // map_T_scan (the first scan: identity)
  sr_result r{};
};

static void odometry_check_scan(const sr_xyz_view* scan) {
  validate_view(scan);
  if (scan->n == 0) fail(SR_ERR_INVALID_ARGUMENT, "GICP requires non-empty source and target point clouds");
  if (scan->n > 0xFFFFFFF0ull) fail(SR_ERR_INVALID_ARGUMENT, "point clouds above 4G points are not supported");
}

// Voxel filter of the scan and, from the second scan on, its grid and normals.
static void odometry_prepare(sr_odometry* o, const sr_xyz_view& scan, OdoScan& sc, hipStream_t s) {
  Receipt rc;
  Cloud raw;
  pack_cloud(scan.x, scan.y, scan.z, uint32_t(scan.n), scan.mem == SR_MEM_DEVICE, raw, s, &rc.h2d);
  voxel_downsample(raw, o->cfg.voxel_leaf, nullptr, false, sc.vox, s);
  std::memset(&sc.r, 0, sizeof(sc.r));
  if (o->scans > 0) prepare_cloud(sc.vox.centroids, o->gicp, sc.src, rc, s, false, o->cfg.voxel_leaf);
}

// Registers each scan against its odometry's map in one batch, from the
// constant-velocity guess (last pose composed with the last scan-to-scan
// motion). Fails, naming the scan, before any map changes.
static void odometry_register(const std::vector<sr_odometry*>& o, const std::vector<OdoScan*>& sc,
                              Batch& B, hipStream_t s) {
  const size_t n = o.size();
  std::vector<const Prepared*> src(n), tgt(n);
  std::vector<sr_gicp_config> cfgs(n);
  std::vector<double> pose7(7 * n);
  for (size_t i = 0; i < n; ++i) {
    const Pose64 guess = compose64(o[i]->pose, compose64(inverse64(o[i]->prev), o[i]->pose));
    for (int a = 0; a < 4; ++a) pose7[7 * i + a] = guess.q[a];
    for (int a = 0; a < 3; ++a) pose7[7 * i + 4 + a] = guess.t[a];
    src[i] = &sc[i]->src;
    tgt[i] = &o[i]->map_prep;
    cfgs[i] = o[i]->gicp;
  }
  batch_setup(src, tgt, cfgs, B, s, pose7.data());
  batch_run(B, s);
  std::vector<InstState> st(n);
  download(st.data(), B.states.p, sizeof(InstState) * n, s);
  for (size_t i = 0; i < n; ++i) {
    std::string msg;
    const int32_t code = state_to_result(st[i], o[i]->gicp, sc[i]->src, o[i]->map_prep, &sc[i]->r, &msg);
    if (code) fail(code, "scan " + std::to_string(o[i]->scans) + ": " + msg);
    for (int a = 0; a < 4; ++a) sc[i]->pose.q[a] = st[i].q[a];
    for (int a = 0; a < 3; ++a) sc[i]->pose.t[a] = st[i].t[a];
  }
}

// Merges the registered scan into the map and advances the odometry.
static void odometry_commit(sr_odometry* o, OdoScan& sc, hipStream_t s) {
  sr_result& r = sc.r;
  if (o->scans == 0) {
    r.q[3] = 1.0f;
    r.q64[3] = 1.0;
    r.converged = 1;
    r.source_points = sc.vox.cells;
  }
  MapPose mp;
  quat_mat64(sc.pose.q, mp.r);
  for (int a = 0; a < 3; ++a) mp.t[a] = sc.pose.t[a];
  Receipt rc;
  const uint64_t drops = o->map.drops;
  const uint32_t n_old = o->map.list.n;
  o->map.insert(sc.vox.centroids.pts.p, sc.vox.cells, mp, s);
  odometry_update_map(o, o->map.drops == drops ? n_old : 0u, rc, s);
  o->prev = o->scans > 0 ? o->pose : sc.pose;
  o->pose = sc.pose;
  ++o->scans;
  r.target_points = o->map.list.n;
}

extern "C" {

int32_t sr_odometry_create(const sr_odometry_config* config, const sr_gicp_config* gicp,
                           sr_odometry** out) {
  return guarded([&] {
    if (!out) fail(SR_ERR_INVALID_ARGUMENT, "out must not be null");
    *out = nullptr;
    if (!config || !gicp) fail(SR_ERR_INVALID_ARGUMENT, "odometry config must not be null");
    if (!(config->voxel_leaf > 0.0f)) fail(SR_ERR_INVALID_ARGUMENT, "leaf_size must be greater than zero");
    if (config->max_map_points == 0 || config->max_map_points > 0x7FFFFFFFull)
      fail(SR_ERR_INVALID_ARGUMENT, "max_map_points must be in [1, 2^31)");
    check_k(gicp->k_neighbors, gicp->covariance_radius);
    ensure_device();
    auto o = std::make_unique<sr_odometry>();
    o->cfg = *config;
    o->gicp = *gicp;
    o->gicp.voxel_leaf = 0.0f;  // scans are filtered here, the map holds voxels already
    hipStream_t s = thread_stream();
    o->map.init(config->voxel_leaf, uint32_t(config->max_map_points), s);
    SR_HIP(hipStreamSynchronize(s));
    *out = o.release();
  });
}

void sr_odometry_destroy(sr_odometry* odo) {
  if (!odo) return;
  (void)hipDeviceSynchronize();
  delete odo;
}

int32_t sr_odometry_push(sr_odometry* odo, const sr_xyz_view* scan, sr_result* out) {
  return guarded([&] {
    if (!odo) fail(SR_ERR_INVALID_ARGUMENT, "odometry must not be null");
    odometry_check_scan(scan);
    hipStream_t s = thread_stream();
    OdoScan sc;
    odometry_prepare(odo, *scan, sc, s);
    if (odo->scans > 0) {
      Batch B;
      ProfScope prof("odometry_register", s);
      odometry_register({odo}, {&sc}, B, s);
    }
    ProfScope prof_map("odometry_map", s);
    odometry_commit(odo, sc, s);
    if (out) *out = sc.r;
  });
}

int32_t sr_odometry_push_many(sr_odometry* const* odos, const sr_xyz_view* scans, uint32_t count,
                              sr_result* out) {
  return guarded([&] {
    if (count == 0) return;
    if (!odos || !scans) fail(SR_ERR_INVALID_ARGUMENT, "odometry arguments must not be null");
    for (uint32_t i = 0; i < count; ++i) {
      if (!odos[i]) fail(SR_ERR_INVALID_ARGUMENT, "odometry must not be null");
      for (uint32_t j = 0; j < i; ++j)
        if (odos[j] == odos[i]) fail(SR_ERR_INVALID_ARGUMENT, "an odometry appears twice in one call");
      odometry_check_scan(&scans[i]);
    }
    ensure_device();
    std::vector<OdoScan> sc(count);
    // scans filtered and prepared on the workers, registered in one batch,
    // merged into their maps on the workers
    Workers::get().run(count, [&](size_t i, hipStream_t s) { odometry_prepare(odos[i], scans[i], sc[i], s); });
    std::vector<sr_odometry*> reg;
    std::vector<OdoScan*> reg_sc;
    for (uint32_t i = 0; i < count; ++i)
      if (odos[i]->scans > 0) {
        reg.push_back(odos[i]);
        reg_sc.push_back(&sc[i]);
      }
    if (!reg.empty()) {
      hipStream_t s = thread_stream();
      Batch B;
      ProfScope prof("odometry_register", s);
      odometry_register(reg, reg_sc, B, s);
    }
    Workers::get().run(count, [&](size_t i, hipStream_t s) { odometry_commit(odos[i], sc[i], s); });
    if (out)
      for (uint32_t i = 0; i < count; ++i) out[i] = sc[i].r;
  });
}

uint64_t sr_odometry_map_points(const sr_odometry* odo) { return odo ? odo->map.list.n : 0; }

int32_t sr_odometry_map_stats(const sr_odometry* odo, uint64_t* out4) {
  return guarded([&] {
    if (!odo || !out4) fail(SR_ERR_INVALID_ARGUMENT, "null odometry argument");
    out4[0] = odo->map.list.n;
    out4[1] = odo->refreshed;
    out4[2] = odo->rebuilds;
    out4[3] = odo->scans;
  });
}

int32_t sr_odometry_map_read(const sr_odometry* odo, float* xyz, uint64_t cap, uint64_t* n) {
  return guarded([&] {
    if (!odo || !n) fail(SR_ERR_INVALID_ARGUMENT, "null odometry argument");
    *n = odo->map.list.n;
    const uint64_t m = std::min<uint64_t>(cap, odo->map.list.n);
    if (!m) return;
    if (!xyz) fail(SR_ERR_INVALID_ARGUMENT, "xyz must not be null");
    std::vector<float4> host(m);
    hipStream_t s = thread_stream();
    download(host.data(), odo->map.list.pts.p, sizeof(float4) * m, s);
    for (uint64_t i = 0; i < m; ++i) {
      xyz[3 * i] = host[i].x;
      xyz[3 * i + 1] = host[i].y;
      xyz[3 * i + 2] = host[i].z;
    }
  });
}

}  // extern "C"
