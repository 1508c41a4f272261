// Sort-free voxel-centroid downsample on gfx950, plus the small primitives it
// needs (packing, bounds, exclusive scan).
//
// Semantics follow VoxelGridDownsample's centroid path for a plain xyz cloud
// (crates/spatialrust-filtering/src/voxel.rs:251-368, 550-636, 1067-1121):
//  - origin: the per-axis minimum (default) or an explicit origin;
//  - key per axis: (v - origin) * inv_leaf truncated to u32 (default origin, when
//    the span fits u32), truncated to i64 (default origin otherwise) or floored
//    to i64 (explicit origin); inv_leaf = 1/leaf computed in f32 on the host;
//  - cells in first-occurrence order (the reference's hash assigns ids on first
//    insert);
//  - centroid = (f32 sequential sum over the cell's points in input order) *
//    (1/count).
// The reference builds cells with a serial hash map; here every point inserts
// its 64-bit linear key into an open-addressing table with atomics, the first
// occurrence of a cell is the point holding the slot's minimum index, and an
// exclusive scan over first-occurrence flags gives each cell its output rank.
// Keys and occupancy are bit-exact; centroids are bit-exact too because each
// cell's members are put back in index order before the f32 sum.

#include "internal.hpp"

namespace sr {

// ---------------------------------------------------------------- utils ----

__global__ void k_pack(const float* __restrict__ x, const float* __restrict__ y,
                       const float* __restrict__ z, uint32_t n, float4* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float4(x[i], y[i], z[i], 0.0f);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

static unsigned fill_blocks(size_t n) {
  size_t b = (n + kBlock - 1) / kBlock;
  return unsigned(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks(n)), dim3(kBlock), 0, s, p, v, n);
}

__global__ void k_bounds_init(uint32_t* b) {
  if (threadIdx.x < 3) b[threadIdx.x] = 0xFF800000u;        // f2ord(+inf)
  else if (threadIdx.x < 6) b[threadIdx.x] = 0x007FFFFFu;   // f2ord(-inf)
}

// Per-axis min/max; NaN coordinates are ignored, like f32::min/max. Each block
// reduces its stride of points in registers, across the wave by shuffles and
// across waves in LDS, then issues one atomic per axis (few blocks, so the
// atomics on the six words never queue up).
__global__ void k_bounds(const float4* __restrict__ pts, uint32_t n, uint32_t* __restrict__ b) {
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float4 p = pts[i];
    mn[0] = fminf(mn[0], p.x);
    mn[1] = fminf(mn[1], p.y);
    mn[2] = fminf(mn[2], p.z);
    mx[0] = fmaxf(mx[0], p.x);
    mx[1] = fmaxf(mx[1], p.y);
    mx[2] = fmaxf(mx[2], p.z);
  }
  for (int off = kWave / 2; off > 0; off >>= 1) {
    for (int a = 0; a < 3; ++a) {
      mn[a] = fminf(mn[a], __shfl_xor(mn[a], off));
      mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off));
    }
  }
  __shared__ float smn[kBlock / kWave][3], smx[kBlock / kWave][3];
  const int wid = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    for (int a = 0; a < 3; ++a) {
      smn[wid][a] = mn[a];
      smx[wid][a] = mx[a];
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int a = threadIdx.x;
    float lo = smn[0][a], hi = smx[0][a];
    for (int w = 1; w < kBlock / kWave; ++w) {
      lo = fminf(lo, smn[w][a]);
      hi = fmaxf(hi, smx[w][a]);
    }
    if (lo <= hi) {
      atomicMin(&b[a], f2ord(lo));
      atomicMax(&b[3 + a], f2ord(hi));
    }
  }
}

void cloud_bounds(const float4* pts, uint32_t n, uint32_t* dev_bounds, hipStream_t s) {
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, s, dev_bounds);
  if (n == 0) return;
  unsigned blocks = unsigned(std::min<size_t>((n + kBlock * 8 - 1) / (kBlock * 8), 512));
  hipLaunchKernelGGL(k_bounds, dim3(blocks), dim3(kBlock), 0, s, pts, n, dev_bounds);
}

void pack_cloud(const float* x, const float* y, const float* z, uint32_t n, bool device_mem,
                Cloud& out, hipStream_t s, uint64_t* h2d_bytes) {
  out.n = n;
  out.pts.resize(n > 0 ? n : 1);
  if (n == 0) return;
  unsigned blocks = (n + kBlock - 1) / kBlock;
  if (device_mem) {
    hipLaunchKernelGGL(k_pack, dim3(blocks), dim3(kBlock), 0, s, x, y, z, n, out.pts.p);
    return;
  }
  DevBuf<float> tmp(size_t(3) * n);
  SR_HIP(hipMemcpyAsync(tmp.p, x, sizeof(float) * n, hipMemcpyHostToDevice, s));
  SR_HIP(hipMemcpyAsync(tmp.p + n, y, sizeof(float) * n, hipMemcpyHostToDevice, s));
  SR_HIP(hipMemcpyAsync(tmp.p + 2 * size_t(n), z, sizeof(float) * n, hipMemcpyHostToDevice, s));
  if (h2d_bytes) *h2d_bytes += uint64_t(12) * n;
  hipLaunchKernelGGL(k_pack, dim3(blocks), dim3(kBlock), 0, s, tmp.p, tmp.p + n, tmp.p + 2 * size_t(n),
                     n, out.pts.p);
  // tmp returns to the caching allocator; the copy and the pack are ordered on s,
  // and any later reuse of the block is ordered after them on the same stream.
  SR_HIP(hipStreamSynchronize(s));
}

// ----------------------------------------------------------------- scan ----

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;  // 2048

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wave_tot,
                                                         uint32_t& block_total) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  uint32_t incl = v;
  for (int off = 1; off < kWave; off <<= 1) {
    uint32_t u = __shfl_up(incl, off);
    if (lane >= off) incl += u;
  }
  if (lane == kWave - 1) wave_tot[wid] = incl;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
  for (int w = 0; w < kBlock / kWave; ++w) {
    uint32_t t = wave_tot[w];
    if (w < wid) pre += t;
    tot += t;
  }
  block_total = tot;
  __syncthreads();
  return pre + incl - v;
}

// Scans one tile of kScanTile values in place in LDS-staged form.
__global__ void k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                             uint32_t n, uint32_t* __restrict__ tile_sums,
                             const uint32_t* __restrict__ tile_offsets, uint32_t* total) {
  __shared__ uint32_t tile[kScanTile];
  __shared__ uint32_t wave_tot[kBlock / kWave];
  const size_t base = size_t(blockIdx.x) * kScanTile;
  for (int k = 0; k < kScanItems; ++k) {
    size_t i = base + size_t(k) * kBlock + threadIdx.x;
    tile[k * kBlock + threadIdx.x] = i < n ? in[i] : 0u;
  }
  __syncthreads();
  uint32_t loc[kScanItems], sum = 0;
  for (int k = 0; k < kScanItems; ++k) {
    loc[k] = tile[threadIdx.x * kScanItems + k];
    sum += loc[k];
  }
  uint32_t btot;
  uint32_t pre = block_exclusive_scan(sum, wave_tot, btot);
  if (tile_offsets) pre += tile_offsets[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    tile[threadIdx.x * kScanItems + k] = pre;
    pre += loc[k];
  }
  __syncthreads();
  if (out) {
    for (int k = 0; k < kScanItems; ++k) {
      size_t i = base + size_t(k) * kBlock + threadIdx.x;
      if (i < n) out[i] = tile[k * kBlock + threadIdx.x];
    }
  }
  if (threadIdx.x == 0) {
    if (tile_sums) tile_sums[blockIdx.x] = btot;
    if (total && gridDim.x == 1) *total = btot;
  }
}

// Sums of each tile only (first level of the multi-tile scan).
__global__ void k_tile_sums(const uint32_t* __restrict__ in, uint32_t n,
                            uint32_t* __restrict__ tile_sums) {
  __shared__ uint32_t wave_tot[kBlock / kWave];
  const size_t base = size_t(blockIdx.x) * kScanTile;
  uint32_t sum = 0;
  for (int k = 0; k < kScanItems; ++k) {
    size_t i = base + size_t(k) * kBlock + threadIdx.x;
    if (i < n) sum += in[i];
  }
  uint32_t btot;
  (void)block_exclusive_scan(sum, wave_tot, btot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = btot;
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total,
                        hipStream_t s) {
  if (n == 0) {
    if (total) fill_u32(total, 0u, 1, s);
    return;
  }
  uint32_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles == 1) {
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kBlock), 0, s, in, out, n, nullptr, nullptr,
                       total);
    return;
  }
  DevBuf<uint32_t> sums(tiles);
  hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(kBlock), 0, s, in, n, sums.p);
  exclusive_scan_u32(sums.p, sums.p, tiles, total, s);
  hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kBlock), 0, s, in, out, n, nullptr, sums.p,
                     nullptr);
}

// ---------------------------------------------------------------- voxel ----

struct VoxelParams {
  float org[3];
  float inv_leaf;
  int32_t mode;   // 0: u32 trunc from min, 1: i64 trunc from min, 2: floor from origin
  int32_t error;  // 1: linear key range does not fit 64 bits
  int64_t kmin[3];
  unsigned long long dims[3];
};

__device__ __forceinline__ int64_t f32_to_i64_sat(float v) {
  if (!(v == v)) return 0;
  if (v >= 9.2233720368547758e18f) return INT64_MAX;
  if (v <= -9.2233720368547758e18f) return INT64_MIN;
  return (int64_t)v;
}

__device__ __forceinline__ uint32_t f32_to_u32_sat(float v) {
  if (!(v == v) || v <= 0.0f) return 0u;
  if (v >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)v;
}

__device__ __forceinline__ void voxel_key(const VoxelParams& P, float x, float y, float z,
                                          int64_t k[3]) {
  float f[3] = {__fmul_rn(__fsub_rn(x, P.org[0]), P.inv_leaf),
                __fmul_rn(__fsub_rn(y, P.org[1]), P.inv_leaf),
                __fmul_rn(__fsub_rn(z, P.org[2]), P.inv_leaf)};
  for (int a = 0; a < 3; ++a) {
    if (P.mode == 0) k[a] = int64_t(f32_to_u32_sat(f[a]));
    else if (P.mode == 1) k[a] = f32_to_i64_sat(f[a]);
    else k[a] = f32_to_i64_sat(floorf(f[a]));
  }
}

// Origin, key mode and key ranges from the bounds.
__device__ void voxel_params_from_bounds(const uint32_t* __restrict__ b, float leaf_inv,
                                         int explicit_origin, float o0, float o1, float o2,
                                         VoxelParams* P) {
  float mn[3] = {ord2f(b[0]), ord2f(b[1]), ord2f(b[2])};
  float mx[3] = {ord2f(b[3]), ord2f(b[4]), ord2f(b[5])};
  VoxelParams p;
  p.inv_leaf = leaf_inv;
  p.error = 0;
  if (explicit_origin) {
    p.org[0] = o0;
    p.org[1] = o1;
    p.org[2] = o2;
    p.mode = 2;
  } else {
    for (int a = 0; a < 3; ++a) p.org[a] = mn[a];
    const float limit = 4294967296.0f;  // fits_u32_voxel_key, voxel.rs:1081-1086
    bool fits = __fmul_rn(__fsub_rn(mx[0], mn[0]), leaf_inv) <= limit &&
                __fmul_rn(__fsub_rn(mx[1], mn[1]), leaf_inv) <= limit &&
                __fmul_rn(__fsub_rn(mx[2], mn[2]), leaf_inv) <= limit;
    p.mode = fits ? 0 : 1;
  }
  int64_t klo[3], khi[3];
  voxel_key(p, mn[0], mn[1], mn[2], klo);
  voxel_key(p, mx[0], mx[1], mx[2], khi);
  double prod = 1.0;
  for (int a = 0; a < 3; ++a) {
    p.kmin[a] = klo[a];
    long long span = khi[a] - klo[a];
    p.dims[a] = (unsigned long long)(span < 0 ? 0 : span) + 1ull;
    prod *= double(p.dims[a]);
  }
  if (!(prod < 1.8e19)) p.error = 1;
  *P = p;
}

__global__ void k_voxel_params(const uint32_t* __restrict__ b, float leaf_inv, int explicit_origin,
                               float o0, float o1, float o2, VoxelParams* P) {
  if (threadIdx.x == 0) voxel_params_from_bounds(b, leaf_inv, explicit_origin, o0, o1, o2, P);
}

__device__ __forceinline__ uint32_t hash64(unsigned long long k, uint32_t mask) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return uint32_t(k) & mask;
}

constexpr unsigned long long kEmptyKey = ~0ull;

__device__ __forceinline__ unsigned long long linear_voxel_key(const VoxelParams& P, const float4& p) {
  int64_t k[3];
  voxel_key(P, p.x, p.y, p.z, k);
  return ((unsigned long long)(k[2] - P.kmin[2]) * P.dims[1] +
          (unsigned long long)(k[1] - P.kmin[1])) * P.dims[0] +
         (unsigned long long)(k[0] - P.kmin[0]);
}

__global__ void k_voxel_insert(const float4* __restrict__ pts, uint32_t n,
                               const VoxelParams* __restrict__ Pp,
                               unsigned long long* __restrict__ tkeys, uint32_t* __restrict__ tfirst,
                               uint32_t mask,
                               uint32_t* __restrict__ slot_of) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const VoxelParams P = *Pp;
  if (P.error) {  // the call fails; keep the later kernels' indices in bounds
    slot_of[i] = 0;
    return;
  }
  float4 p = pts[i];
  const unsigned long long lin = linear_voxel_key(P, p);
  uint32_t h = hash64(lin, mask);
  while (true) {
    // A slot's key never changes once set, so a plain read that finds this
    // key (or another one) decides without the atomic; a stale empty read
    // only falls through to it.
    const unsigned long long cur = tkeys[h];
    if (cur == lin) break;
    if (cur == kEmptyKey) {
      const unsigned long long prev = atomicCAS(&tkeys[h], kEmptyKey, lin);
      if (prev == kEmptyKey) {
        tfirst[h] = i;  // a first guess at the slot's lowest index (k_voxel_first_fix)
        break;
      }
      if (prev == lin) break;
    }
    h = (h + 1) & mask;
  }
  slot_of[i] = h;
}

// The point that claimed a slot wrote its index as the slot's first; the
// slot's lower indices (points of the same voxel that came later) lower it.
// Most voxels hold one point, or were claimed by their lowest index, so few
// points take the atomic.
__global__ void k_voxel_first_fix(const uint32_t* __restrict__ slot_of, uint32_t n,
                                  uint32_t* __restrict__ tfirst) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t* f = &tfirst[slot_of[i]];
  if (*f > i) atomicMin(f, i);
}

// First-occurrence flags; also clears the per-cell counters (cells <= points)
// and the four list counters clear4[0..3] of the centroid kernels.
__global__ void k_voxel_flags(const uint32_t* __restrict__ slot_of,
                              const uint32_t* __restrict__ tfirst, uint32_t n,
                              uint32_t* __restrict__ flag, uint32_t* __restrict__ ocount,
                              uint32_t* __restrict__ clear4) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 4) clear4[i] = 0u;
  if (i < n) {
    flag[i] = tfirst[slot_of[i]] == i ? 1u : 0u;
    ocount[i] = 0u;
  }
}

__global__ void k_voxel_cellinfo(const float4* __restrict__ pts, const uint32_t* __restrict__ slot_of,
                                 const uint32_t* __restrict__ flag, const uint32_t* __restrict__ rank,
                                 uint32_t n, const VoxelParams* __restrict__ Pp,
                                 uint32_t* __restrict__ cell_of_slot, int64_t* __restrict__ okeys) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  uint32_t c = rank[i], s = slot_of[i];
  cell_of_slot[s] = c;
  if (okeys) {
    int64_t k[3];
    float4 p = pts[i];
    voxel_key(*Pp, p.x, p.y, p.z, k);
    okeys[3 * size_t(c) + 0] = k[0];
    okeys[3 * size_t(c) + 1] = k[1];
    okeys[3 * size_t(c) + 2] = k[2];
  }
}

// Points per output cell. The cells are numbered densely in first-occurrence
// order, so these counters are a compact array (the hash table's slots are
// spread over twice the points and their atomics cost several times more).
// Each point keeps its cell and its arrival rank in the cell (the counter's
// old value), so the scatter below needs no second round of atomics.
__global__ void k_voxel_count(const uint32_t* __restrict__ slot_of,
                              const uint32_t* __restrict__ cell_of_slot, uint32_t n,
                              uint32_t* __restrict__ ocount, uint32_t* __restrict__ cell_of,
                              uint32_t* __restrict__ rank_in_cell) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // (the clamp only matters when a key-range error has left slots unnumbered)
  const uint32_t c = min(cell_of_slot[slot_of[i]], n - 1);
  cell_of[i] = c;
  rank_in_cell[i] = atomicAdd(&ocount[c], 1u);
}

// Members of each cell in arrival order (the centroid kernels put them in
// index order).
__global__ void k_voxel_scatter(const uint32_t* __restrict__ cell_of,
                                const uint32_t* __restrict__ rank_in_cell,
                                const uint32_t* __restrict__ cstart, uint32_t n,
                                uint32_t* __restrict__ members) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  members[min(cstart[cell_of[i]] + rank_in_cell[i], n - 1)] = i;
}

constexpr uint32_t kSmallCell = 64;

// Restores index order inside a segment (insertion sort; segments are short).
__device__ __forceinline__ void sort_segment(uint32_t* seg, uint32_t cnt) {
  for (uint32_t a = 1; a < cnt; ++a) {
    uint32_t v = seg[a];
    uint32_t b = a;
    while (b > 0 && seg[b - 1] > v) {
      seg[b] = seg[b - 1];
      --b;
    }
    seg[b] = v;
  }
}

__device__ __forceinline__ float4 cell_centroid(const float4* __restrict__ pts,
                                                const uint32_t* seg, uint32_t cnt) {
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for (uint32_t a = 0; a < cnt; ++a) {
    float4 p = pts[seg[a]];
    sx = __fadd_rn(sx, p.x);
    sy = __fadd_rn(sy, p.y);
    sz = __fadd_rn(sz, p.z);
  }
  float inv = __fdiv_rn(1.0f, (float)cnt);
  return make_float4(__fmul_rn(sx, inv), __fmul_rn(sy, inv), __fmul_rn(sz, inv), 0.0f);
}

// Cells of at most kTinyCell members are sorted and summed by one thread here,
// their member indices sorted in registers by a network of the next power of
// two; up to kSmallCell members by one wave (k_voxel_medium_cells); larger
// ones by a workgroup (k_voxel_big_cells). (A per-thread insertion sort in
// device memory took a 64-member cell ~1000 dependent memory round trips.)
constexpr uint32_t kTinyCell = 16;

template <int N>
__device__ __forceinline__ float4 small_cell_centroid(const float4* __restrict__ pts,
                                                      const uint32_t* __restrict__ seg,
                                                      uint32_t cnt) {
  uint32_t v[N];
#pragma unroll
  for (int t = 0; t < N; ++t) v[t] = uint32_t(t) < cnt ? seg[t] : 0xFFFFFFFFu;
  sort_net<N>(v);
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;  // sequential sums in index order
#pragma unroll
  for (int t = 0; t < N; ++t)
    if (uint32_t(t) < cnt) {
      const float4 q = pts[v[t]];
      sx = __fadd_rn(sx, q.x);
      sy = __fadd_rn(sy, q.y);
      sz = __fadd_rn(sz, q.z);
    }
  const float inv = __fdiv_rn(1.0f, (float)cnt);
  return make_float4(__fmul_rn(sx, inv), __fmul_rn(sy, inv), __fmul_rn(sz, inv), 0.0f);
}

__global__ void k_voxel_centroids(const float4* __restrict__ pts, const uint32_t* __restrict__ ncells,
                                  const uint32_t* __restrict__ cstart,
                                  const uint32_t* __restrict__ ocount, uint32_t* __restrict__ members,
                                  float4* __restrict__ out, uint32_t* __restrict__ big_list,
                                  uint32_t* __restrict__ big_count,
                                  uint32_t* __restrict__ med_count) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nc = *ncells;
  const uint32_t cnt = c < nc ? ocount[c] : 0u;
  // list appends aggregated per wave: one atomic per wave and list
  const bool big = cnt > kSmallCell, med = cnt > kTinyCell && !big;
  const int lane = int(threadIdx.x & 63u);
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long mb = __ballot(big), mm = __ballot(med);
  if (mb) {
    const int leader = __ffsll((long long)mb) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(big_count, uint32_t(__popcll(mb)));
    base = __shfl(base, leader);
    if (big) big_list[base + uint32_t(__popcll(mb & below))] = c;
  }
  if (mm) {  // medium list: filled from the end of the same buffer
    const int leader = __ffsll((long long)mm) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(med_count, uint32_t(__popcll(mm)));
    base = __shfl(base, leader);
    if (med) big_list[nc - 1u - (base + uint32_t(__popcll(mm & below)))] = c;
  }
  if (c >= nc || big || med) return;
  const uint32_t* seg = members + cstart[c];
  float4 r;
  if (cnt <= 4) r = small_cell_centroid<4>(pts, seg, cnt);
  else if (cnt <= 8) r = small_cell_centroid<8>(pts, seg, cnt);
  else r = small_cell_centroid<16>(pts, seg, cnt);
  out[c] = r;
}

// One wave per medium cell (kTinyCell < members <= kSmallCell = 64): the
// member indices, one per lane, are put in ascending order by a bitonic
// network across the lanes, the points loaded in parallel, and their
// coordinates added in index order (lane reads), which is the sequential f32
// sum cell_centroid forms.
static_assert(kSmallCell <= 64, "one member per lane");
__global__ void __launch_bounds__(kBlock) k_voxel_medium_cells(
    const float4* __restrict__ pts, const uint32_t* __restrict__ ncells,
    const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ ocount,
    const uint32_t* __restrict__ members, const uint32_t* __restrict__ list,
    const uint32_t* __restrict__ med_count, float4* __restrict__ out) {
  const uint32_t nmed = *med_count, nc = *ncells;
  const int lane = int(threadIdx.x & 63u);
  const uint32_t waves = gridDim.x * (blockDim.x / 64u);
  for (uint32_t w = blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u; w < nmed; w += waves) {
    const uint32_t c = list[nc - 1u - w];
    const uint32_t cnt = ocount[c];
    uint32_t v = uint32_t(lane) < cnt ? members[cstart[c] + uint32_t(lane)] : 0xFFFFFFFFu;
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        const uint32_t o = uint32_t(__shfl_xor(int(v), stride));
        const bool up = (lane & size) == 0, lower = (lane & stride) == 0;
        // the lower lane of a pair keeps the smaller key in an ascending run
        v = (lower == up) ? min(v, o) : max(v, o);
      }
    }
    float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (uint32_t(lane) < cnt) p = pts[v];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (uint32_t a = 0; a < cnt; ++a) {
      sx = __fadd_rn(sx, __shfl(p.x, int(a)));
      sy = __fadd_rn(sy, __shfl(p.y, int(a)));
      sz = __fadd_rn(sz, __shfl(p.z, int(a)));
    }
    if (lane == 0) {
      const float inv = __fdiv_rn(1.0f, (float)cnt);
      out[c] = make_float4(__fmul_rn(sx, inv), __fmul_rn(sy, inv), __fmul_rn(sz, inv), 0.0f);
    }
  }
}

// Large cells: one workgroup per cell sorts its member indices (bitonic in LDS
// for up to kBigTile members, selection in global memory beyond) then one lane
// sums in index order.
constexpr uint32_t kBigTile = 4096;

__global__ void k_voxel_big_cells(const float4* __restrict__ pts,
                                  const uint32_t* __restrict__ cstart,
                                  const uint32_t* __restrict__ ocount, uint32_t* __restrict__ members,
                                  const uint32_t* __restrict__ big_list,
                                  const uint32_t* __restrict__ big_count, float4* __restrict__ out) {
  __shared__ uint32_t keys[kBigTile];
  const uint32_t nbig = *big_count;
  for (uint32_t b = blockIdx.x; b < nbig; b += gridDim.x) {
    const uint32_t c = big_list[b];
    const uint32_t cnt = ocount[c];
    uint32_t* seg = members + cstart[c];
    if (cnt <= kBigTile) {
      uint32_t m = 1;
      while (m < cnt) m <<= 1;
      for (uint32_t a = threadIdx.x; a < m; a += blockDim.x) keys[a] = a < cnt ? seg[a] : 0xFFFFFFFFu;
      __syncthreads();
      for (uint32_t size = 2; size <= m; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
          for (uint32_t a = threadIdx.x; a < m; a += blockDim.x) {
            uint32_t partner = a ^ stride;
            if (partner > a) {
              bool up = (a & size) == 0;
              uint32_t ka = keys[a], kb = keys[partner];
              if ((ka > kb) == up) {
                keys[a] = kb;
                keys[partner] = ka;
              }
            }
          }
          __syncthreads();
        }
      }
      for (uint32_t a = threadIdx.x; a < cnt; a += blockDim.x) seg[a] = keys[a];
      __syncthreads();
    } else if (threadIdx.x == 0) {
      sort_segment(seg, cnt);
    }
    __syncthreads();
    if (threadIdx.x == 0) out[c] = cell_centroid(pts, seg, cnt);
    __syncthreads();
  }
}

__global__ void k_error_check(const VoxelParams* P, uint32_t* flag_out) {
  if (threadIdx.x == 0) *flag_out = uint32_t(P->error);
}

void voxel_downsample(const Cloud& in, float leaf, const float* origin, bool want_keys,
                      VoxelOut& out, hipStream_t s) {
  const uint32_t n = in.n;
  out.cells = 0;
  if (n == 0) return;
  const float inv_leaf = 1.0f / leaf;  // voxel.rs:267, f32 on the host
  DevBuf<uint32_t> bounds(8);
  cloud_bounds(in.pts.p, n, bounds.p, s);
  DevBuf<VoxelParams> params(1);
  hipLaunchKernelGGL(k_voxel_params, dim3(1), dim3(64), 0, s, bounds.p, inv_leaf,
                     origin ? 1 : 0, origin ? origin[0] : 0.0f, origin ? origin[1] : 0.0f,
                     origin ? origin[2] : 0.0f, params.p);
  uint32_t cap = 1024;
  while (cap < 2u * n) cap <<= 1;
  DevBuf<unsigned long long> tkeys(cap);
  DevBuf<uint32_t> tfirst(cap), cell_of_slot(cap);
  hipLaunchKernelGGL(k_fill_u64, dim3(fill_blocks(cap)), dim3(kBlock), 0, s, tkeys.p, kEmptyKey,
                     size_t(cap));
  DevBuf<uint32_t> slot_of(n), flag(n), rank(n), ocount(n), cstart(n), cursor(n), members(n),
      scal(5);
  const unsigned blocks = (n + kBlock - 1) / kBlock;
  ProfScope prof_insert("voxel_insert", s);
  hipLaunchKernelGGL(k_voxel_insert, dim3(blocks), dim3(kBlock), 0, s, in.pts.p, n, params.p,
                     tkeys.p, tfirst.p, cap - 1, slot_of.p);
  hipLaunchKernelGGL(k_voxel_first_fix, dim3(blocks), dim3(kBlock), 0, s, slot_of.p, n, tfirst.p);
  hipLaunchKernelGGL(k_voxel_flags, dim3(blocks), dim3(kBlock), 0, s, slot_of.p, tfirst.p, n, flag.p,
                     ocount.p, scal.p + 1);  // (scal[1..4]: list counters and error word)
  exclusive_scan_u32(flag.p, rank.p, n, scal.p, s);  // scal[0] = cell count
  if (want_keys) out.keys.resize(size_t(3) * n);
  hipLaunchKernelGGL(k_voxel_cellinfo, dim3(blocks), dim3(kBlock), 0, s, in.pts.p, slot_of.p,
                     flag.p, rank.p, n, params.p, cell_of_slot.p, want_keys ? out.keys.p : nullptr);
  // flag[] is free after k_voxel_cellinfo: it takes each point's cell
  hipLaunchKernelGGL(k_voxel_count, dim3(blocks), dim3(kBlock), 0, s, slot_of.p, cell_of_slot.p, n,
                     ocount.p, flag.p, cursor.p);
  exclusive_scan_u32(ocount.p, cstart.p, n, nullptr, s);
  hipLaunchKernelGGL(k_voxel_scatter, dim3(blocks), dim3(kBlock), 0, s, flag.p, cursor.p, cstart.p, n,
                     members.p);
  out.centroids.pts.resize(n);
  DevBuf<uint32_t> big_list(n);
  hipLaunchKernelGGL(k_voxel_centroids, dim3(blocks), dim3(kBlock), 0, s, in.pts.p, scal.p,
                     cstart.p, ocount.p, members.p, out.centroids.pts.p, big_list.p, scal.p + 1,
                     scal.p + 4);
  hipLaunchKernelGGL(k_voxel_medium_cells, dim3(1024), dim3(kBlock), 0, s, in.pts.p, scal.p,
                     cstart.p, ocount.p, members.p, big_list.p, scal.p + 4, out.centroids.pts.p);
  hipLaunchKernelGGL(k_voxel_big_cells, dim3(256), dim3(kBlock), 0, s, in.pts.p, cstart.p,
                     ocount.p, members.p, big_list.p, scal.p + 1, out.centroids.pts.p);
  hipLaunchKernelGGL(k_error_check, dim3(1), dim3(64), 0, s, params.p, scal.p + 2);
  uint32_t host[3];
  download(host, scal.p, sizeof(host), s);
  if (host[2]) fail(3, "voxel key range does not fit a 64-bit linear key; use a larger leaf_size");
  out.cells = host[0];
  out.centroids.n = host[0];
  out.counts = std::move(ocount);
}

// ---- many clouds in one set of launches -----------------------------------
//
// The clouds are concatenated (segment s = points [off[s], off[s+1])). Every
// per-point kernel finds its segment from its workgroup (a segment's points
// start a fresh workgroup), so a segment's results are those of a launch over
// it alone: each has its own bounds, key parameters and hash-table region, and
// the global first-occurrence order keeps the segments apart and each in its
// own first-occurrence order.

constexpr uint32_t kSegBoundsChunk = kBlock * 8;  // points per workgroup of k_seg_bounds

// Workgroup min/max of per-thread partial bounds, merged into the segment's
// six ordered words (one atomic per word per workgroup).
__device__ __forceinline__ void merge_block_bounds(float (&mn)[3], float (&mx)[3],
                                                   uint32_t* __restrict__ b6) {
  for (int o = kWave / 2; o > 0; o >>= 1) {
    for (int a = 0; a < 3; ++a) {
      mn[a] = fminf(mn[a], __shfl_xor(mn[a], o));
      mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], o));
    }
  }
  __shared__ float smn[kBlock / kWave][3], smx[kBlock / kWave][3];
  const int wid = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    for (int a = 0; a < 3; ++a) {
      smn[wid][a] = mn[a];
      smx[wid][a] = mx[a];
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int a = threadIdx.x;
    float l = smn[0][a], h = smx[0][a];
    for (int w = 1; w < kBlock / kWave; ++w) {
      l = fminf(l, smn[w][a]);
      h = fmaxf(h, smx[w][a]);
    }
    if (l <= h) {
      atomicMin(&b6[a], f2ord(l));
      atomicMax(&b6[3 + a], f2ord(h));
    }
  }
}

// Workgroup w of segment s packs points [w*kSegBoundsChunk, ...) of it and,
// when b is set, merges their bounds into the segment's words (b holds
// f2ord(+inf) / f2ord(-inf) on entry).
__global__ void k_seg_pack(const SegSource* __restrict__ src, uint32_t nseg,
                           const uint32_t* __restrict__ blk_begin, float4* __restrict__ out,
                           uint32_t* __restrict__ b) {
  const uint32_t s = seg_of(blk_begin, nseg, blockIdx.x);
  const SegSource S = src[s];
  const uint32_t lo = (blockIdx.x - blk_begin[s]) * kSegBoundsChunk;
  const uint32_t hi = min(S.n, lo + kSegBoundsChunk);
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const float x = S.x[i], y = S.y[i], z = S.z[i];
    out[S.offset + i] = make_float4(x, y, z, 0.0f);
    mn[0] = fminf(mn[0], x);
    mn[1] = fminf(mn[1], y);
    mn[2] = fminf(mn[2], z);
    mx[0] = fmaxf(mx[0], x);
    mx[1] = fmaxf(mx[1], y);
    mx[2] = fmaxf(mx[2], z);
  }
  if (b) merge_block_bounds(mn, mx, b + 6 * s);
}

__global__ void k_seg_bounds_init(uint32_t* __restrict__ b, uint32_t nseg) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 6 * nseg) b[t] = (t % 6) < 3 ? 0xFF800000u : 0x007FFFFFu;  // f2ord(+inf) / f2ord(-inf)
}

// k_bounds per segment: workgroup w of segment s reduces points
// [off[s] + w*kSegBoundsChunk, ...) of it.
__global__ void k_seg_bounds(const float4* __restrict__ pts, const uint32_t* __restrict__ off,
                             uint32_t nseg, const uint32_t* __restrict__ blk_begin,
                             uint32_t* __restrict__ b) {
  const uint32_t s = seg_of(blk_begin, nseg, blockIdx.x);
  const uint32_t lo = off[s] + (blockIdx.x - blk_begin[s]) * kSegBoundsChunk;
  const uint32_t hi = min(off[s + 1], lo + kSegBoundsChunk);
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    float4 p = pts[i];
    mn[0] = fminf(mn[0], p.x);
    mn[1] = fminf(mn[1], p.y);
    mn[2] = fminf(mn[2], p.z);
    mx[0] = fmaxf(mx[0], p.x);
    mx[1] = fmaxf(mx[1], p.y);
    mx[2] = fmaxf(mx[2], p.z);
  }
  merge_block_bounds(mn, mx, b + 6 * s);
}

__global__ void k_seg_voxel_params(const uint32_t* __restrict__ b, uint32_t nseg, float leaf_inv,
                                   VoxelParams* __restrict__ P, uint32_t* __restrict__ err_word) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) *err_word = 0u;  // the key-range error flag of k_seg_voxel_tail
  if (s < nseg) voxel_params_from_bounds(b + 6 * s, leaf_inv, 0, 0.0f, 0.0f, 0.0f, P + s);
}

struct SegTable {
  uint32_t base;  // first slot of the segment's hash-table region
  uint32_t mask;  // region size - 1 (a power of two)
};

__global__ void k_seg_voxel_insert(const float4* __restrict__ pts, const uint32_t* __restrict__ off,
                                   uint32_t nseg, const uint32_t* __restrict__ blk_begin,
                                   const VoxelParams* __restrict__ Pp,
                                   const SegTable* __restrict__ tab,
                                   unsigned long long* __restrict__ tkeys,
                                   uint32_t* __restrict__ tfirst,
                                   uint32_t* __restrict__ slot_of) {
  const uint32_t s = seg_of(blk_begin, nseg, blockIdx.x);
  const uint32_t i = off[s] + (blockIdx.x - blk_begin[s]) * blockDim.x + threadIdx.x;
  if (i >= off[s + 1]) return;
  const VoxelParams P = Pp[s];
  if (P.error) {  // the call fails; keep the later kernels' indices in bounds
    slot_of[i] = 0;
    return;
  }
  const SegTable T = tab[s];
  const unsigned long long lin = linear_voxel_key(P, pts[i]);
  uint32_t h = hash64(lin, T.mask);
  while (true) {
    // A slot's key never changes once set, so a plain read that finds this
    // key (or another one) decides without the atomic; a stale empty read
    // only falls through to it.
    const unsigned long long cur = tkeys[T.base + h];
    if (cur == lin) break;
    if (cur == kEmptyKey) {
      const unsigned long long prev = atomicCAS(&tkeys[T.base + h], kEmptyKey, lin);
      if (prev == kEmptyKey) {
        tfirst[T.base + h] = i;  // a first guess at the slot's lowest index (k_voxel_first_fix)
        break;
      }
      if (prev == lin) break;
    }
    h = (h + 1) & T.mask;
  }
  slot_of[i] = T.base + h;
}

// Per segment: its first output cell (the exclusive rank of its first point),
// and a key-range error of any segment.
__global__ void k_seg_voxel_tail(const uint32_t* __restrict__ rank, const uint32_t* __restrict__ off,
                                 uint32_t nseg, const VoxelParams* __restrict__ P,
                                 const uint32_t* __restrict__ total, uint32_t* __restrict__ out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nseg) {
    out[s] = off[s] < off[nseg] ? rank[off[s]] : *total;
    if (P[s].error) atomicOr(&out[nseg + 1], 1u);
  }
  if (s == 0) out[nseg] = *total;
}

// Workgroups per segment, as a launch's first-workgroup table.
static std::vector<uint32_t> seg_blocks(const std::vector<uint32_t>& off, uint32_t per_block) {
  std::vector<uint32_t> b(off.size());
  b[0] = 0;
  for (size_t s = 0; s + 1 < off.size(); ++s)
    b[s + 1] = b[s] + std::max<uint32_t>(1, (off[s + 1] - off[s] + per_block - 1) / per_block);
  return b;
}

void pack_many(const std::vector<SegSource>& srcs, Cloud& out, hipStream_t s,
               DevBuf<uint8_t>* hold, const uint32_t** dev_bounds) {
  std::vector<uint32_t> off(srcs.size() + 1, 0);
  for (size_t i = 0; i < srcs.size(); ++i) off[i + 1] = off[i] + srcs[i].n;
  out.n = off.back();
  out.pts.resize(std::max<uint32_t>(out.n, 1));
  if (srcs.empty() || out.n == 0) return;
  std::vector<uint32_t> blk = seg_blocks(off, kSegBoundsChunk);
  // the bounds words start as f2ord(+inf) / f2ord(-inf) and go up with the
  // tables (one copy for all three)
  std::vector<uint32_t> binit(hold ? 6 * srcs.size() : 0);
  for (size_t t = 0; t < binit.size(); ++t) binit[t] = (t % 6) < 3 ? 0xFF800000u : 0x007FFFFFu;
  HostPack pk;
  const size_t o_src = pk.add(srcs), o_blk = pk.add(blk), o_b = pk.add(binit);
  DevBuf<uint8_t> tables(pk.bytes.size());
  upload(tables.p, pk.bytes.data(), pk.bytes.size(), s);
  uint32_t* b = hold ? reinterpret_cast<uint32_t*>(tables.p + o_b) : nullptr;
  hipLaunchKernelGGL(k_seg_pack, dim3(blk.back()), dim3(kBlock), 0, s,
                     HostPack::at<SegSource>(tables.p, o_src), uint32_t(srcs.size()),
                     HostPack::at<uint32_t>(tables.p, o_blk), out.pts.p, b);
  if (hold) {  // the caller keeps the tables (and the bounds in them)
    *dev_bounds = b;
    *hold = std::move(tables);
    return;
  }
  // the tables outlive the kernel: released at the next wait on s
  if (!tables.release_after(s)) SR_HIP(hipStreamSynchronize(s));
}

void seg_bounds(const float4* pts, const std::vector<uint32_t>& off, uint32_t* dev_bounds,
                hipStream_t s) {
  const uint32_t nseg = uint32_t(off.size() - 1);
  std::vector<uint32_t> blk = seg_blocks(off, kSegBoundsChunk);
  HostPack pk;
  const size_t o_off = pk.add(off), o_blk = pk.add(blk);
  DevBuf<uint8_t> tables(pk.bytes.size());
  upload(tables.p, pk.bytes.data(), pk.bytes.size(), s);
  hipLaunchKernelGGL(k_seg_bounds_init, dim3((6 * nseg + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     dev_bounds, nseg);
  hipLaunchKernelGGL(k_seg_bounds, dim3(blk.back()), dim3(kBlock), 0, s, pts,
                     HostPack::at<uint32_t>(tables.p, o_off), nseg,
                     HostPack::at<uint32_t>(tables.p, o_blk), dev_bounds);
  if (!tables.release_after(s)) SR_HIP(hipStreamSynchronize(s));
}

void voxel_downsample_many(const Cloud& in, const std::vector<uint32_t>& off, float leaf,
                           VoxelOut& out, std::vector<uint32_t>& cell_off, hipStream_t s,
                           const uint32_t* dev_bounds) {
  const uint32_t n = in.n, nseg = uint32_t(off.size() - 1);
  out.cells = 0;
  cell_off.assign(nseg + 1, 0);
  if (n == 0 || nseg == 0) return;
  const float inv_leaf = 1.0f / leaf;  // voxel.rs:267, f32 on the host
  DevBuf<uint32_t> bounds;
  if (!dev_bounds) {
    bounds.resize(6 * size_t(nseg));
    seg_bounds(in.pts.p, off, bounds.p, s);
    dev_bounds = bounds.p;
  }
  DevBuf<VoxelParams> params(nseg);
  DevBuf<uint32_t> tailv(nseg + 2);  // per-segment first cells, the total, the error flag
  hipLaunchKernelGGL(k_seg_voxel_params, dim3((nseg + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     dev_bounds, nseg, inv_leaf, params.p, tailv.p + nseg + 1);
  // one hash-table region per segment, sized as voxel_downsample sizes its table
  std::vector<SegTable> tab(nseg);
  uint64_t slots = 0;
  for (uint32_t g = 0; g < nseg; ++g) {
    uint32_t cap = 1024;
    while (cap < 2u * (off[g + 1] - off[g])) cap <<= 1;
    tab[g] = SegTable{uint32_t(slots), cap - 1};
    slots += cap;
  }
  if (slots > 0xFFFFFFFFull) fail(3, "too many points for one batch");
  const uint32_t cap = uint32_t(slots);
  std::vector<uint32_t> blk = seg_blocks(off, kBlock);
  HostPack pk;
  const size_t o_tab = pk.add(tab), o_off = pk.add(off), o_blk = pk.add(blk);
  DevBuf<uint8_t> tables(pk.bytes.size());
  upload(tables.p, pk.bytes.data(), pk.bytes.size(), s);  // one copy for the three tables
  const SegTable* d_tab = HostPack::at<SegTable>(tables.p, o_tab);
  const uint32_t* d_off = HostPack::at<uint32_t>(tables.p, o_off);
  const uint32_t* d_blk = HostPack::at<uint32_t>(tables.p, o_blk);
  DevBuf<unsigned long long> tkeys(cap);
  DevBuf<uint32_t> tfirst(cap), cell_of_slot(cap);
  hipLaunchKernelGGL(k_fill_u64, dim3(fill_blocks(cap)), dim3(kBlock), 0, s, tkeys.p, kEmptyKey,
                     size_t(cap));
  DevBuf<uint32_t> slot_of(n), flag(n), rank(n), ocount(n), cstart(n), cursor(n), members(n),
      scal(5);
  const unsigned blocks = (n + kBlock - 1) / kBlock;
  ProfScope prof_insert("voxel_insert", s);
  hipLaunchKernelGGL(k_seg_voxel_insert, dim3(blk.back()), dim3(kBlock), 0, s, in.pts.p, d_off,
                     nseg, d_blk, params.p, d_tab, tkeys.p, tfirst.p, slot_of.p);
  hipLaunchKernelGGL(k_voxel_first_fix, dim3(blocks), dim3(kBlock), 0, s, slot_of.p, n, tfirst.p);
  hipLaunchKernelGGL(k_voxel_flags, dim3(blocks), dim3(kBlock), 0, s, slot_of.p, tfirst.p, n, flag.p,
                     ocount.p, scal.p + 1);  // (scal[1..4]: list counters and error word)
  exclusive_scan_u32(flag.p, rank.p, n, scal.p, s);  // scal[0] = cell count
  hipLaunchKernelGGL(k_voxel_cellinfo, dim3(blocks), dim3(kBlock), 0, s, in.pts.p, slot_of.p,
                     flag.p, rank.p, n, params.p, cell_of_slot.p, (int64_t*)nullptr);
  // flag[] is free after k_voxel_cellinfo: it takes each point's cell
  hipLaunchKernelGGL(k_voxel_count, dim3(blocks), dim3(kBlock), 0, s, slot_of.p, cell_of_slot.p, n,
                     ocount.p, flag.p, cursor.p);
  exclusive_scan_u32(ocount.p, cstart.p, n, nullptr, s);
  hipLaunchKernelGGL(k_voxel_scatter, dim3(blocks), dim3(kBlock), 0, s, flag.p, cursor.p, cstart.p, n,
                     members.p);
  out.centroids.pts.resize(n);
  DevBuf<uint32_t> big_list(n);
  hipLaunchKernelGGL(k_voxel_centroids, dim3(blocks), dim3(kBlock), 0, s, in.pts.p, scal.p,
                     cstart.p, ocount.p, members.p, out.centroids.pts.p, big_list.p, scal.p + 1,
                     scal.p + 4);
  hipLaunchKernelGGL(k_voxel_medium_cells, dim3(1024), dim3(kBlock), 0, s, in.pts.p, scal.p,
                     cstart.p, ocount.p, members.p, big_list.p, scal.p + 4, out.centroids.pts.p);
  hipLaunchKernelGGL(k_voxel_big_cells, dim3(256), dim3(kBlock), 0, s, in.pts.p, cstart.p,
                     ocount.p, members.p, big_list.p, scal.p + 1, out.centroids.pts.p);
  hipLaunchKernelGGL(k_seg_voxel_tail, dim3((nseg + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     rank.p, d_off, nseg, params.p, scal.p, tailv.p);
  std::vector<uint32_t> host(nseg + 2);
  download(host.data(), tailv.p, sizeof(uint32_t) * host.size(), s);
  if (host[nseg + 1]) fail(3, "voxel key range does not fit a 64-bit linear key; use a larger leaf_size");
  for (uint32_t g = 0; g <= nseg; ++g) cell_off[g] = host[g];
  out.cells = host[nseg];
  out.centroids.n = host[nseg];
  out.counts = std::move(ocount);
}

}  // namespace sr
