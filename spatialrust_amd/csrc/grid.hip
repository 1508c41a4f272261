// Dense uniform grid over a device cloud: counting sort into CSR cells, points
// re-laid in grid order as float4 (x, y, z, original index bits) so a cell scan
// is one contiguous 16-byte-per-point read.
//
// Layout and clamping follow grid_bounds/build_grid
// (crates/spatialrust-search/src/uniform_grid.rs:9-87): origin = per-axis min,
// dims = floor(span * inv_cell) + 1, cell id (cz*dy + cy)*dx + cx, points in a
// cell ordered by ascending original index. The reference builds the grid on
// the CPU and serially; here counting, scan and scatter run on the device and
// each cell's members are put back in index order afterwards.

#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace sr {

constexpr uint64_t kMaxCells = 64000000ull;  // MAX_UNIFORM_GRID_CELLS, uniform_grid.rs:6

void host_bounds(const Cloud& c, uint32_t* host6, hipStream_t s) {
  DevBuf<uint32_t> b(8);
  cloud_bounds(c.pts.p, c.n, b.p, s);
  SR_HIP(hipMemcpyAsync(host6, b.p, 6 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  SR_HIP(hipStreamSynchronize(s));
}

static void grid_desc(const uint32_t* hb, float h, uint32_t npoints, GridDesc& d) {
  float mn[3] = {ord2f(hb[0]), ord2f(hb[1]), ord2f(hb[2])};
  float mx[3] = {ord2f(hb[3]), ord2f(hb[4]), ord2f(hb[5])};
  d.ox = mn[0];
  d.oy = mn[1];
  d.oz = mn[2];
  d.h = h;
  d.inv_h = 1.0f / h;
  int32_t dims[3];
  float extent = 0.0f;
  for (int a = 0; a < 3; ++a) {
    float span = std::floor((mx[a] - mn[a]) * d.inv_h);
    if (!(span >= 0.0f)) span = 0.0f;
    dims[a] = int32_t(std::min(span, 2.0e9f)) + 1;
    extent = std::max(extent, std::fabs(mn[a]) + std::fabs(mx[a]));
  }
  d.dx = dims[0];
  d.dy = dims[1];
  d.dz = dims[2];
  uint64_t cells = uint64_t(dims[0]) * uint64_t(dims[1]) * uint64_t(dims[2]);
  d.ncells = uint32_t(std::min<uint64_t>(cells, 0xFFFFFFFFull));
  d.npoints = npoints;
  // Filing slack: f32 rounding of (v - o) * inv_h is ~1.2e-7 relative to the
  // cell coordinate, plus the rounding of inv_h itself, plus the rounding of the
  // box corners o + c*h relative to the coordinate magnitudes.
  int32_t dmax = std::max(dims[0], std::max(dims[1], dims[2]));
  d.margin = h * std::max(1e-4f, float(dmax) * 1e-6f) + extent * 1e-6f;
  d.wpr = (uint32_t(d.dx) + 31) / 32;
  d.bits = nullptr;
  d.wps = (uint32_t(d.dy) + 31) / 32;
  d.rows = nullptr;
  d.pad2_ = 0;
}

__global__ void k_grid_count(const float4* __restrict__ pts, uint32_t n, GridDesc g,
                             uint32_t* __restrict__ counts, uint32_t* __restrict__ cid,
                             uint32_t* __restrict__ nonempty) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool first = false;
  if (i < n) {
    float4 p = pts[i];
    int cx = cell_coord(p.x, g.ox, g.inv_h, g.dx);
    int cy = cell_coord(p.y, g.oy, g.inv_h, g.dy);
    int cz = cell_coord(p.z, g.oz, g.inv_h, g.dz);
    uint32_t c = (uint32_t(cz) * uint32_t(g.dy) + uint32_t(cy)) * uint32_t(g.dx) + uint32_t(cx);
    uint32_t old = atomicAdd(&counts[c], 1u);
    if (cid) cid[i] = c;
    first = old == 0;
  }
  if (nonempty) {
    // one atomic per wave instead of one per newly touched cell
    const unsigned long long m = __ballot(first);
    if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicAdd(nonempty, uint32_t(__popcll(m)));
  }
}

__global__ void k_grid_scatter(const uint32_t* __restrict__ cid, uint32_t n,
                               const uint32_t* __restrict__ start, uint32_t* __restrict__ cursor,
                               uint32_t* __restrict__ order) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t c = cid[i];
  order[start[c] + atomicAdd(&cursor[c], 1u)] = i;
}

constexpr uint32_t kSmallSeg = 48;
constexpr uint32_t kBigSeg = 4096;

__global__ void k_sort_cells(const uint32_t* __restrict__ start, uint32_t ncells,
                             uint32_t* __restrict__ order, uint32_t* __restrict__ big_list,
                             uint32_t* __restrict__ big_count) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  uint32_t b = start[c], e = start[c + 1];
  uint32_t cnt = e - b;
  if (cnt < 2) return;
  if (cnt > kSmallSeg) {
    big_list[atomicAdd(big_count, 1u)] = c;
    return;
  }
  uint32_t* seg = order + b;
  for (uint32_t a = 1; a < cnt; ++a) {
    uint32_t v = seg[a];
    uint32_t k = a;
    while (k > 0 && seg[k - 1] > v) {
      seg[k] = seg[k - 1];
      --k;
    }
    seg[k] = v;
  }
}

__global__ void k_sort_big_cells(const uint32_t* __restrict__ start, uint32_t* __restrict__ order,
                                 const uint32_t* __restrict__ big_list,
                                 const uint32_t* __restrict__ big_count) {
  __shared__ uint32_t keys[kBigSeg];
  const uint32_t nbig = *big_count;
  for (uint32_t bi = blockIdx.x; bi < nbig; bi += gridDim.x) {
    uint32_t c = big_list[bi];
    uint32_t b = start[c], cnt = start[c + 1] - b;
    uint32_t* seg = order + b;
    if (cnt <= kBigSeg) {
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
    } else if (threadIdx.x == 0) {
      // Oversized cell: shell sort in global memory (rare; keeps exact order).
      for (uint32_t gap = cnt / 2; gap > 0; gap /= 2) {
        for (uint32_t a = gap; a < cnt; ++a) {
          uint32_t v = seg[a];
          uint32_t k = a;
          while (k >= gap && seg[k - gap] > v) {
            seg[k] = seg[k - gap];
            k -= gap;
          }
          seg[k] = v;
        }
      }
    }
    __syncthreads();
  }
}

// Occupancy bitmap: one thread per 32-cell word of an x-row.
__global__ void k_grid_bits(const uint32_t* __restrict__ cs, GridDesc g, uint32_t nwords,
                            uint32_t* __restrict__ bits) {
  const uint32_t wi = blockIdx.x * blockDim.x + threadIdx.x;
  if (wi >= nwords) return;
  const uint32_t row = wi / g.wpr, w = wi % g.wpr;
  const uint32_t base = row * uint32_t(g.dx);
  uint32_t m = 0;
  for (uint32_t b = 0; b < 32; ++b) {
    const uint32_t cx = w * 32 + b;
    if (cx >= uint32_t(g.dx)) break;
    if (cs[base + cx + 1] > cs[base + cx]) m |= 1u << b;
  }
  bits[wi] = m;
}

// Row bitmap: one thread per 32-row word of a z-slab; bit set when any cell of
// the x-row is occupied.
__global__ void k_grid_rowbits(const uint32_t* __restrict__ bits, GridDesc g, uint32_t nwords,
                               uint32_t* __restrict__ rows) {
  const uint32_t wi = blockIdx.x * blockDim.x + threadIdx.x;
  if (wi >= nwords) return;
  const uint32_t cz = wi / g.wps, wy = wi % g.wps;
  uint32_t m = 0;
  for (uint32_t b = 0; b < 32; ++b) {
    const uint32_t cy = wy * 32 + b;
    if (cy >= uint32_t(g.dy)) break;
    const uint32_t* rb = bits + (size_t(cz) * uint32_t(g.dy) + cy) * g.wpr;
    uint32_t any = 0;
    for (uint32_t w = 0; w < g.wpr; ++w) any |= rb[w];
    if (any) m |= 1u << b;
  }
  rows[wi] = m;
}

__global__ void k_grid_gather(const float4* __restrict__ pts, const uint32_t* __restrict__ order,
                              uint32_t n, float4* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t o = order[i];
  float4 p = pts[o];
  out[i] = make_float4(p.x, p.y, p.z, __uint_as_float(o));
}

__global__ void k_gather_to_grid(const float4* __restrict__ orig, const uint32_t* __restrict__ order,
                                 uint32_t n, float4* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = orig[order[i]];
}

void gather_to_grid_order(const Grid& g, const float4* orig, float4* out, hipStream_t s) {
  uint32_t n = g.desc.npoints;
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_to_grid, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, orig,
                     g.order.p, n, out);
}

// Picks the cell edge from the measured mean occupancy of non-empty cells:
// surface clouds fill cells in proportion to h^2, so one correction from a
// first guess lands close to the requested occupancy.
float choose_cell_edge(const Cloud& c, const uint32_t* hb, float occupancy, float leaf_hint) {
  hipStream_t s = thread_stream();
  float span[3];
  for (int a = 0; a < 3; ++a) span[a] = std::max(ord2f(hb[3 + a]) - ord2f(hb[a]), 0.0f);
  float smax = std::max(span[0], std::max(span[1], span[2]));
  if (c.n <= 1 || !(smax > 0.0f)) return smax > 0.0f ? smax : 1.0f;
  // First guess: points spread over the bounding box's largest face.
  float s2 = 1.0f;
  {
    float a[3] = {span[0], span[1], span[2]};
    std::sort(a, a + 3);
    s2 = std::max(a[1] * a[2], smax * smax * 1e-6f);
  }
  float h = std::sqrt(s2 * occupancy / float(c.n));
  if (leaf_hint > 0.0f) {
    // One point per voxel of edge leaf_hint on surfaces: a cell of edge h
    // holds about (h / leaf)^2 of them.
    h = leaf_hint * std::sqrt(occupancy);
  }
  DevBuf<uint32_t> nonempty(1);
  for (int pass = 0; pass < 2; ++pass) {
    // keep the dense grid under the cell cap
    for (int guard = 0; guard < 64; ++guard) {
      GridDesc d;
      grid_desc(hb, h, c.n, d);
      uint64_t cells = uint64_t(d.dx) * uint64_t(d.dy) * uint64_t(d.dz);
      if (cells <= kMaxCells) break;
      h *= 1.25f;
    }
    GridDesc d;
    grid_desc(hb, h, c.n, d);
    DevBuf<uint32_t> counts(d.ncells);
    fill_u32(counts.p, 0u, d.ncells, s);
    fill_u32(nonempty.p, 0u, 1, s);
    hipLaunchKernelGGL(k_grid_count, dim3((c.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       c.pts.p, c.n, d, counts.p, (uint32_t*)nullptr, nonempty.p);
    uint32_t ne = 0;
    SR_HIP(hipMemcpyAsync(&ne, nonempty.p, 4, hipMemcpyDeviceToHost, s));
    SR_HIP(hipStreamSynchronize(s));
    float occ = float(c.n) / float(std::max(ne, 1u));
    float scale = std::sqrt(occupancy / occ);
    scale = std::min(std::max(scale, 0.25f), 4.0f);
    if (std::fabs(scale - 1.0f) < 0.15f) break;
    h *= scale;
  }
  for (int guard = 0; guard < 64; ++guard) {
    GridDesc d;
    grid_desc(hb, h, c.n, d);
    if (uint64_t(d.dx) * uint64_t(d.dy) * uint64_t(d.dz) <= kMaxCells) break;
    h *= 1.25f;
  }
  return h;
}

void build_grid(const Cloud& c, const uint32_t* hb, float h, Grid& g, hipStream_t s) {
  grid_desc(hb, h, c.n, g.desc);
  uint64_t cells = uint64_t(g.desc.dx) * uint64_t(g.desc.dy) * uint64_t(g.desc.dz);
  if (cells > kMaxCells) {
    fail(3, "grid would need " + std::to_string(cells) + " cells (cap " +
                std::to_string(kMaxCells) + "); use a larger radius or the CPU path");
  }
  const uint32_t n = c.n, nc = g.desc.ncells;
  g.cell_start.resize(size_t(nc) + 1);
  g.pts.resize(n > 0 ? n : 1);
  g.order.resize(n > 0 ? n : 1);
  DevBuf<uint32_t> counts(size_t(nc) + 1), cid(n > 0 ? n : 1), cursor(nc), scal(2);
  fill_u32(counts.p, 0u, size_t(nc) + 1, s);
  fill_u32(cursor.p, 0u, nc, s);
  if (n > 0) {
    hipLaunchKernelGGL(k_grid_count, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, c.pts.p, n,
                       g.desc, counts.p, cid.p, (uint32_t*)nullptr);
  }
  exclusive_scan_u32(counts.p, g.cell_start.p, nc + 1, nullptr, s);
  {
    g.desc.wpr = (uint32_t(g.desc.dx) + 31) / 32;
    const uint64_t nwords = uint64_t(g.desc.wpr) * uint64_t(g.desc.dy) * uint64_t(g.desc.dz);
    g.occ.resize(std::max<uint64_t>(nwords, 1));
    g.desc.bits = g.occ.p;
    hipLaunchKernelGGL(k_grid_bits, dim3(uint32_t((nwords + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       g.cell_start.p, g.desc, uint32_t(nwords), g.occ.p);
    g.desc.wps = (uint32_t(g.desc.dy) + 31) / 32;
    const uint64_t nrw = uint64_t(g.desc.wps) * uint64_t(g.desc.dz);
    g.rowocc.resize(std::max<uint64_t>(nrw, 1));
    g.desc.rows = g.rowocc.p;
    hipLaunchKernelGGL(k_grid_rowbits, dim3(uint32_t((nrw + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       g.occ.p, g.desc, uint32_t(nrw), g.rowocc.p);
  }
  if (n == 0) return;
  hipLaunchKernelGGL(k_grid_scatter, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, cid.p, n,
                     g.cell_start.p, cursor.p, g.order.p);
  fill_u32(scal.p, 0u, 1, s);
  DevBuf<uint32_t> big(std::max<uint32_t>(nc, 1));
  hipLaunchKernelGGL(k_sort_cells, dim3((nc + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     g.cell_start.p, nc, g.order.p, big.p, scal.p);
  hipLaunchKernelGGL(k_sort_big_cells, dim3(128), dim3(kBlock), 0, s, g.cell_start.p, g.order.p,
                     big.p, scal.p);
  hipLaunchKernelGGL(k_grid_gather, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, c.pts.p,
                     g.order.p, n, g.pts.p);
  // counts/cid/cursor/big return to the cache; ordering on s keeps reuse safe.
}

// Chebyshev distance transform of occupancy, separable: along one axis
// d'(c) = min_k max(|k|, d(c + k e_axis)) for |k| <= cap; three passes give the
// min over the cube. Values saturate at cap + 1 ("at least cap + 1").
__global__ void k_empty_init(const uint32_t* __restrict__ cs, uint32_t ncells,
                             uint8_t* __restrict__ out) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) out[c] = cs[c + 1] > cs[c] ? 0 : uint8_t(kEmptyCap + 1);
}

__global__ void k_empty_pass(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, GridDesc g,
                             int axis) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncells) return;
  const uint32_t dx = uint32_t(g.dx), dy = uint32_t(g.dy);
  const int cx = int(c % dx), cy = int((c / dx) % dy), cz = int(c / (dx * dy));
  const int coord = axis == 0 ? cx : (axis == 1 ? cy : cz);
  const int dim = axis == 0 ? g.dx : (axis == 1 ? g.dy : g.dz);
  const int stride = axis == 0 ? 1 : (axis == 1 ? g.dx : g.dx * g.dy);
  int best = in[c];
  for (int k = 1; k <= kEmptyCap && k < best; ++k) {
    if (coord - k >= 0) best = min(best, max(k, int(in[int64_t(c) - int64_t(k) * stride])));
    if (coord + k < dim) best = min(best, max(k, int(in[int64_t(c) + int64_t(k) * stride])));
  }
  out[c] = uint8_t(best);
}

void build_empty_radius(Grid& g, hipStream_t s) {
  const uint32_t nc = g.desc.ncells;
  g.empty_radius.resize(std::max<uint32_t>(nc, 1));
  if (nc == 0) return;
  DevBuf<uint8_t> tmp(nc);
  const dim3 grid((nc + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_empty_init, grid, dim3(kBlock), 0, s, g.cell_start.p, nc, g.empty_radius.p);
  hipLaunchKernelGGL(k_empty_pass, grid, dim3(kBlock), 0, s, g.empty_radius.p, tmp.p, g.desc, 0);
  hipLaunchKernelGGL(k_empty_pass, grid, dim3(kBlock), 0, s, tmp.p, g.empty_radius.p, g.desc, 1);
  hipLaunchKernelGGL(k_empty_pass, grid, dim3(kBlock), 0, s, g.empty_radius.p, tmp.p, g.desc, 2);
  SR_HIP(hipMemcpyAsync(g.empty_radius.p, tmp.p, nc, hipMemcpyDeviceToDevice, s));
}

}  // namespace sr
