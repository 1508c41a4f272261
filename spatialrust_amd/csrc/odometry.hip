// Rolling voxel map for scan-to-map odometry.
//
// The map is a list of voxel representatives (float4, in insertion order) and
// an open-addressing hash from the voxel key to the list index. A scan, already
// registered, is transformed into the map frame; each of its points whose map
// voxel is not yet present claims it, and of several points of one scan falling
// into the same new voxel the lowest index wins. Winners are appended in point
// order, so the list, and everything built from it, is the same on every run.
// When the list would exceed its capacity the oldest entries are dropped and
// the hash is rebuilt from the survivors.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sr_gicp.h"
#include "internal.hpp"

namespace sr {

namespace {

constexpr int kBlk = 256;
constexpr uint64_t kEmptyKey = ~0ull;

__device__ __forceinline__ uint64_t voxel_key(const float4& p, float inv_leaf) {
  // floor((p - 0) * inv_leaf) per axis in f32, 21 bits each (offset 2^20)
  const int kx = int(fminf(fmaxf(floorf(p.x * inv_leaf), -1048576.0f), 1048575.0f));
  const int ky = int(fminf(fmaxf(floorf(p.y * inv_leaf), -1048576.0f), 1048575.0f));
  const int kz = int(fminf(fmaxf(floorf(p.z * inv_leaf), -1048576.0f), 1048575.0f));
  return (uint64_t(uint32_t(kx + 1048576)) << 42) | (uint64_t(uint32_t(ky + 1048576)) << 21) |
         uint64_t(uint32_t(kz + 1048576));
}

__device__ __forceinline__ uint32_t key_hash(uint64_t k, uint64_t mask) {
  k ^= k >> 31;
  k *= 0x7FB5D329728EA185ull;
  k ^= k >> 27;
  k *= 0x81DADEF4BC2DD44Dull;
  k ^= k >> 33;
  return uint32_t(k & mask);
}

// p' = R p + t in FP64 (products and sums in a fixed order, no FMA), rounded to f32.
__global__ void k_map_transform(const float4* __restrict__ in, uint32_t n, MapPose P,
                                float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = in[i];
  const double x = p.x, y = p.y, z = p.z;
  const double ox = ((P.r[0] * x + P.r[1] * y) + P.r[2] * z) + P.t[0];
  const double oy = ((P.r[3] * x + P.r[4] * y) + P.r[5] * z) + P.t[1];
  const double oz = ((P.r[6] * x + P.r[7] * y) + P.r[8] * z) + P.t[2];
  out[i] = make_float4(float(ox), float(oy), float(oz), 0.0f);
}

// Finds or claims the slot of each point's voxel; points of voxels absent from
// the map bid for the voxel with their index.
__global__ void k_map_insert(const float4* __restrict__ pts, uint32_t n, float inv_leaf,
                             unsigned long long* __restrict__ keys, const uint32_t* __restrict__ val,
                             uint32_t* __restrict__ claim, uint32_t* __restrict__ slot_of,
                             uint64_t mask, uint32_t* __restrict__ overflow) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t key = voxel_key(pts[i], inv_leaf);
  uint32_t slot = key_hash(key, mask);
  for (uint64_t probe = 0; probe <= mask; ++probe) {
    const unsigned long long k = keys[slot];
    if (k == key) break;
    if (k == kEmptyKey) {
      const unsigned long long prev = atomicCAS(&keys[slot], kEmptyKey, (unsigned long long)key);
      if (prev == kEmptyKey || prev == key) break;
    }
    slot = uint32_t((slot + 1) & mask);
    if (probe == mask) {
      atomicOr(overflow, 1u);
      slot_of[i] = kNone;
      return;
    }
  }
  slot_of[i] = slot;
  if (val[slot] == kNone) atomicMin(&claim[slot], i);
}

__global__ void k_map_flags(const uint32_t* __restrict__ slot_of, uint32_t n,
                            const uint32_t* __restrict__ val, const uint32_t* __restrict__ claim,
                            uint32_t* __restrict__ flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = slot_of[i];
  flag[i] = (s != kNone && val[s] == kNone && claim[s] == i) ? 1u : 0u;
}

__global__ void k_map_append(const float4* __restrict__ pts, const uint32_t* __restrict__ slot_of,
                             const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                             uint32_t n, uint32_t base, float4* __restrict__ list,
                             uint32_t* __restrict__ val, uint32_t* __restrict__ claim) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const uint32_t s = slot_of[i];
  const uint32_t idx = base + pos[i];
  const float4 p = pts[i];
  list[idx] = make_float4(p.x, p.y, p.z, 0.0f);
  val[s] = idx;
  claim[s] = kNone;
}

// Re-inserts the whole list (distinct voxels) into an empty table.
__global__ void k_map_reinsert(const float4* __restrict__ list, uint32_t n, float inv_leaf,
                               unsigned long long* __restrict__ keys, uint32_t* __restrict__ val,
                               uint64_t mask, uint32_t* __restrict__ overflow) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t key = voxel_key(list[i], inv_leaf);
  uint32_t slot = key_hash(key, mask);
  for (uint64_t probe = 0; probe <= mask; ++probe) {
    const unsigned long long prev = atomicCAS(&keys[slot], kEmptyKey, (unsigned long long)key);
    if (prev == kEmptyKey) {
      val[slot] = i;
      return;
    }
    slot = uint32_t((slot + 1) & mask);
  }
  atomicOr(overflow, 1u);
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, size_t n) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

inline uint32_t blocks_for(size_t n) { return uint32_t(std::max<size_t>(1, (n + kBlk - 1) / kBlk)); }

void fill_u64(unsigned long long* p, unsigned long long v, size_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u64, dim3(uint32_t(std::min<size_t>(blocks_for(n), 4096))), dim3(kBlk),
                     0, s, p, v, n);
}

// ---- incremental upkeep of the map's grid and normals ----------------------

// field[cell] = 0 for the cells of the list points [n_old, n) (the new voxels).
__global__ void k_mark_new_cells(const float4* __restrict__ list, uint32_t n_old, uint32_t n,
                                 GridDesc g, uint8_t* __restrict__ field) {
  const uint32_t i = n_old + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = list[i];
  const int cx = cell_coord(p.x, g.ox, g.inv_h, g.dx), cy = cell_coord(p.y, g.oy, g.inv_h, g.dy),
            cz = cell_coord(p.z, g.oz, g.inv_h, g.dz);
  field[(uint32_t(cz) * uint32_t(g.dy) + uint32_t(cy)) * uint32_t(g.dx) + uint32_t(cx)] = 0;
}

// Appends the grid positions whose normal must be computed again. A point in
// a cell at Chebyshev distance m >= 1 (cells) from every new point's cell lies
// more than (m - 1) h - 2 margin from each new point (cells are filed with at
// most `margin` of rounding; m is capped, and the cap is a lower bound); its k
// nearest stay when its k-th neighbour is closer than that (2 more margins
// cover the rounding of the squared distances).
__global__ void k_map_affected(const float4* __restrict__ gp, uint32_t n, GridDesc g,
                               const uint8_t* __restrict__ field, const float* __restrict__ ltau,
                               uint32_t n_old, uint32_t* __restrict__ out,
                               uint32_t* __restrict__ count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool aff = false;
  if (i < n) {
    const float4 p = gp[i];
    const uint32_t orig = __float_as_uint(p.w);
    if (orig >= n_old) {
      aff = true;
    } else {
      const int cx = cell_coord(p.x, g.ox, g.inv_h, g.dx), cy = cell_coord(p.y, g.oy, g.inv_h, g.dy),
                cz = cell_coord(p.z, g.oz, g.inv_h, g.dz);
      const int m = field[(uint32_t(cz) * uint32_t(g.dy) + uint32_t(cy)) * uint32_t(g.dx) + uint32_t(cx)];
      const float lim = float(m - 1) * g.h - 4.0f * g.margin;
      aff = !(lim > 0.0f && ltau[orig] < lim * lim);
    }
  }
  const unsigned long long mask = __ballot(aff);
  if (!mask) return;
  const int lane = int(threadIdx.x & 63u);
  const int leader = __ffsll((long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, uint32_t(__popcll(mask)));
  base = __shfl(base, leader);
  if (aff) out[base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)))] = i;
}

// Grid-order point records from the grid points and the list-order normals.
__global__ void k_gather_normals(const float4* __restrict__ gp, const uint32_t* __restrict__ order,
                                 const float4* __restrict__ lnrm, uint32_t n,
                                 float4* __restrict__ rec) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rec[kRec * size_t(i)] = gp[i];
  rec[kRec * size_t(i) + 1] = lnrm[order[i]];
}

}  // namespace

void map_refresh_set(const Grid& g, const float4* list, uint32_t n_old, const float* ltau,
                     uint32_t* subset, uint32_t* count, hipStream_t s) {
  const GridDesc& d = g.desc;
  const uint32_t n = d.npoints;
  fill_u32(count, 0u, 1, s);
  if (n == 0) return;
  const size_t words = (size_t(d.ncells) + 3) / 4;
  DevBuf<uint8_t> field(4 * words);
  static_assert(kEmptyCap + 1 < 256, "distance field is one byte per cell");
  fill_u32(reinterpret_cast<uint32_t*>(field.p), 0x01010101u * uint32_t(kEmptyCap + 1), words, s);
  if (n > n_old) {
    hipLaunchKernelGGL(k_mark_new_cells, dim3(blocks_for(n - n_old)), dim3(kBlk), 0, s, list, n_old,
                       n, d, field.p);
    chebyshev_field(field.p, d, s);
  }
  hipLaunchKernelGGL(k_map_affected, dim3(blocks_for(n)), dim3(kBlk), 0, s, g.pts.p, n, d, field.p,
                     ltau, n_old, subset, count);
  // the field returns to the cache; ordering on s keeps its reuse safe
}

void gather_list_normals(const Grid& g, const float4* lnrm, float4* rec, hipStream_t s) {
  const uint32_t n = g.desc.npoints;
  if (n)
    hipLaunchKernelGGL(k_gather_normals, dim3(blocks_for(n)), dim3(kBlk), 0, s, g.pts.p, g.order.p,
                       lnrm, n, rec);
}

void VoxelMap::init(float leaf_, uint32_t capacity_, hipStream_t s) {
  leaf = leaf_;
  inv_leaf = 1.0f / leaf_;
  capacity = capacity_;
  uint64_t slots = 1024;
  while (slots < 2ull * capacity_) slots <<= 1;
  mask = slots - 1;
  keys.resize(slots);
  val.resize(slots);
  claim.resize(slots);
  list.pts.resize(std::max<uint32_t>(capacity_, 1));
  list.n = 0;
  flags.resize(1);
  scal.resize(2);
  fill_u64(reinterpret_cast<unsigned long long*>(keys.p), kEmptyKey, slots, s);
  fill_u32(val.p, kNone, slots, s);
  fill_u32(claim.p, kNone, slots, s);
  fill_u32(scal.p, 0u, 2, s);
}

void VoxelMap::drop_oldest(uint32_t count, hipStream_t s) {
  if (count == 0) return;
  ++drops;
  const uint32_t keep = list.n > count ? list.n - count : 0;
  if (keep) {
    DevBuf<float4> tmp(keep);
    SR_HIP(hipMemcpyAsync(tmp.p, list.pts.p + (list.n - keep), sizeof(float4) * keep,
                          hipMemcpyDeviceToDevice, s));
    SR_HIP(hipMemcpyAsync(list.pts.p, tmp.p, sizeof(float4) * keep, hipMemcpyDeviceToDevice, s));
    SR_HIP(hipStreamSynchronize(s));  // tmp is released at scope end
  }
  list.n = keep;
  fill_u64(reinterpret_cast<unsigned long long*>(keys.p), kEmptyKey, mask + 1, s);
  fill_u32(val.p, kNone, mask + 1, s);
  if (keep)
    hipLaunchKernelGGL(k_map_reinsert, dim3(blocks_for(keep)), dim3(kBlk), 0, s, list.pts.p, keep,
                       inv_leaf, reinterpret_cast<unsigned long long*>(keys.p), val.p, mask,
                       scal.p + 1);
}

uint32_t VoxelMap::insert(const float4* scan_local, uint32_t n, const MapPose& pose, hipStream_t s) {
  if (n == 0) return 0;
  DevBuf<float4> world(n);
  DevBuf<uint32_t> slot_of(n), pos(n);
  flags.resize(n);
  hipLaunchKernelGGL(k_map_transform, dim3(blocks_for(n)), dim3(kBlk), 0, s, scan_local, n, pose,
                     world.p);
  uint32_t host[2];
  // Claim pass; while the new voxels do not fit, drop the oldest map voxels
  // (the claims of that attempt are void) and claim again against the rest.
  for (;;) {
    hipLaunchKernelGGL(k_map_insert, dim3(blocks_for(n)), dim3(kBlk), 0, s, world.p, n, inv_leaf,
                       reinterpret_cast<unsigned long long*>(keys.p), val.p, claim.p, slot_of.p,
                       mask, scal.p + 1);
    hipLaunchKernelGGL(k_map_flags, dim3(blocks_for(n)), dim3(kBlk), 0, s, slot_of.p, n, val.p,
                       claim.p, flags.p);
    exclusive_scan_u32(flags.p, pos.p, n, scal.p, s);
    download(host, scal.p, sizeof(host), s);
    if (host[1]) fail(SR_ERR_INVALID_ARGUMENT, "voxel map hash table overflow");
    if (uint64_t(list.n) + host[0] <= capacity) break;
    if (list.n == 0) fail(SR_ERR_INVALID_ARGUMENT, "scan has more new voxels than the map capacity");
    fill_u32(claim.p, kNone, mask + 1, s);
    drop_oldest(uint32_t(std::min<uint64_t>(uint64_t(list.n) + host[0] - capacity, list.n)), s);
  }
  const uint32_t added = host[0];
  hipLaunchKernelGGL(k_map_append, dim3(blocks_for(n)), dim3(kBlk), 0, s, world.p, slot_of.p,
                     flags.p, pos.p, n, list.n, list.pts.p, val.p, claim.p);
  list.n += added;
  return added;
}

}  // namespace sr
