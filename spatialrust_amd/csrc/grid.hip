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
#include <cstring>

#include "internal.hpp"

namespace sr {

constexpr uint64_t kMaxCells = 64000000ull;  // MAX_UNIFORM_GRID_CELLS, uniform_grid.rs:6

void host_bounds(const Cloud& c, uint32_t* host6, hipStream_t s) {
  DevBuf<uint32_t> b(8);
  cloud_bounds(c.pts.p, c.n, b.p, s);
  download(host6, b.p, 6 * sizeof(uint32_t), s);
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
                             uint32_t* __restrict__ rnk) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = pts[i];
  int cx = cell_coord(p.x, g.ox, g.inv_h, g.dx);
  int cy = cell_coord(p.y, g.oy, g.inv_h, g.dy);
  int cz = cell_coord(p.z, g.oz, g.inv_h, g.dz);
  uint32_t c = (uint32_t(cz) * uint32_t(g.dy) + uint32_t(cy)) * uint32_t(g.dx) + uint32_t(cx);
  rnk[i] = atomicAdd(&counts[c], 1u);
  cid[i] = c;
}

// Point i goes to its cell's range at the slot its count increment returned
// (arrival order; k_sort_cells restores index order within each cell). Also
// clears the two list counters clear2[0..1] of the cell sorts.
__global__ void k_grid_scatter(const uint32_t* __restrict__ cid, const uint32_t* __restrict__ rnk,
                               uint32_t n, const uint32_t* __restrict__ start,
                               uint32_t* __restrict__ order, uint32_t* __restrict__ clear2) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2) clear2[i] = 0u;
  if (i >= n) return;
  uint32_t c = cid[i];
  order[start[c] + rnk[i]] = i;
}

constexpr uint32_t kSmallSeg = 16;   // per thread, in registers
constexpr uint32_t kMediumSeg = 64;  // one wave (k_sort_medium_cells)
constexpr uint32_t kBigSeg = 4096;

// A cell of at most N members: loaded into registers (padded with ~0, which
// sorts last), sorted by the network, stored back.
template <int N>
__device__ __forceinline__ void sort_small_cell(uint32_t* seg, uint32_t cnt) {
  uint32_t v[N];
#pragma unroll
  for (int t = 0; t < N; ++t) v[t] = uint32_t(t) < cnt ? seg[t] : 0xFFFFFFFFu;
  sort_net<N>(v);
#pragma unroll
  for (int t = 0; t < N; ++t)
    if (uint32_t(t) < cnt) seg[t] = v[t];
}

// Cells of up to kSmallSeg members are sorted here, in registers; larger ones
// are listed (medium from the end of big_list, big from its start; one atomic
// per wave and list) for k_sort_medium_cells / k_sort_big_cells.
__global__ void k_sort_cells(const uint32_t* __restrict__ start, uint32_t ncells,
                             uint32_t* __restrict__ order, uint32_t* __restrict__ big_list,
                             uint32_t* __restrict__ counts) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t b = c < ncells ? start[c] : 0u, cnt = c < ncells ? start[c + 1] - b : 0u;
  const bool big = cnt > kMediumSeg, med = cnt > kSmallSeg && !big;
  const int lane = int(threadIdx.x & 63u);
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long mb = __ballot(big), mm = __ballot(med);
  if (mb) {
    const int leader = __ffsll((long long)mb) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&counts[0], uint32_t(__popcll(mb)));
    base = __shfl(base, leader);
    if (big) big_list[base + uint32_t(__popcll(mb & below))] = c;
  }
  if (mm) {
    const int leader = __ffsll((long long)mm) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&counts[1], uint32_t(__popcll(mm)));
    base = __shfl(base, leader);
    if (med) big_list[ncells - 1u - (base + uint32_t(__popcll(mm & below)))] = c;
  }
  if (cnt < 2 || big || med) return;
  uint32_t* seg = order + b;
  if (cnt <= 4) sort_small_cell<4>(seg, cnt);
  else if (cnt <= 8) sort_small_cell<8>(seg, cnt);
  else sort_small_cell<16>(seg, cnt);
}

// One wave per listed medium cell: its members, one per lane, sorted by a
// bitonic network across the lanes and written back.
__global__ void __launch_bounds__(kBlock) k_sort_medium_cells(const uint32_t* __restrict__ start,
                                                              uint32_t ncells,
                                                              uint32_t* __restrict__ order,
                                                              const uint32_t* __restrict__ list,
                                                              const uint32_t* __restrict__ counts) {
  const uint32_t nmed = counts[1];
  const int lane = int(threadIdx.x & 63u);
  const uint32_t waves = gridDim.x * (blockDim.x / 64u);
  for (uint32_t w = blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u; w < nmed; w += waves) {
    const uint32_t c = list[ncells - 1u - w];
    const uint32_t b = start[c], cnt = start[c + 1] - b;
    uint32_t v = uint32_t(lane) < cnt ? order[b + uint32_t(lane)] : 0xFFFFFFFFu;
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        const uint32_t o = uint32_t(__shfl_xor(int(v), stride));
        const bool up = (lane & size) == 0, lower = (lane & stride) == 0;
        v = (lower == up) ? min(v, o) : max(v, o);
      }
    }
    if (uint32_t(lane) < cnt) order[b + uint32_t(lane)] = v;
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
__device__ __forceinline__ void grid_bits_word(const uint32_t* __restrict__ cs, const GridDesc& g,
                                               uint32_t wi, uint32_t* __restrict__ bits) {
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

__global__ void k_grid_bits(const uint32_t* __restrict__ cs, GridDesc g, uint32_t nwords,
                            uint32_t* __restrict__ bits) {
  const uint32_t wi = blockIdx.x * blockDim.x + threadIdx.x;
  if (wi < nwords) grid_bits_word(cs, g, wi, bits);
}

// Row bitmap: bit cy of word (cz, cy / 32) is set when any cell of the x-row
// (cy, cz) is occupied. One thread per row (32 per word): each lane ORs its
// row's occupancy words and a ballot forms the word, so a grid of few z-slabs
// still spreads over many threads.
__global__ void k_grid_rowbits(const uint32_t* __restrict__ bits, GridDesc g, uint32_t nwords,
                               uint32_t* __restrict__ rows) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t wi = t >> 5;
  uint32_t any = 0;
  if (wi < nwords) {
    const uint32_t cz = wi / g.wps, cy = (wi % g.wps) * 32u + (t & 31u);
    if (cy < uint32_t(g.dy)) {
      const uint32_t* rb = bits + (size_t(cz) * uint32_t(g.dy) + cy) * g.wpr;
      for (uint32_t w = 0; w < g.wpr; ++w) any |= rb[w];
    }
  }
  const unsigned long long m = __ballot(any != 0u);
  const uint32_t lane = threadIdx.x & 63u;
  if ((lane & 31u) == 0u && wi < nwords) rows[wi] = uint32_t(m >> lane);
}

__global__ void k_grid_gather(const float4* __restrict__ pts, const uint32_t* __restrict__ order,
                              uint32_t n, float4* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t o = order[i];
  float4 p = pts[o];
  out[i] = make_float4(p.x, p.y, p.z, __uint_as_float(o));
}

__global__ void k_gather_records(const float4* __restrict__ gp, const float4* __restrict__ orig,
                                 const uint32_t* __restrict__ order, uint32_t n,
                                 float4* __restrict__ rec) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rec[kRec * size_t(i)] = gp[i];
  rec[kRec * size_t(i) + 1] = orig[order[i]];
}

void gather_records(const Grid& g, const float4* orig, float4* rec, hipStream_t s) {
  uint32_t n = g.desc.npoints;
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_records, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     g.pts.p, orig, g.order.p, n, rec);
}

// The cell edge is picked from the measured mean occupancy of non-empty cells
// at candidate edges around a first guess (sketches below); the host steps are
// shared by the one-cloud and the many-cloud builds, so both pick the same edge
// for the same cloud.
// First guess; returns false when the cloud has no extent (h is then final).
static bool edge_first_guess(const uint32_t* hb, uint32_t n, float occupancy, float leaf_hint,
                             float& h) {
  float span[3];
  for (int a = 0; a < 3; ++a) span[a] = std::max(ord2f(hb[3 + a]) - ord2f(hb[a]), 0.0f);
  float smax = std::max(span[0], std::max(span[1], span[2]));
  if (n <= 1 || !(smax > 0.0f)) {
    h = smax > 0.0f ? smax : 1.0f;
    return false;
  }
  // First guess: points spread over the bounding box's largest face.
  float s2 = 1.0f;
  {
    float a[3] = {span[0], span[1], span[2]};
    std::sort(a, a + 3);
    s2 = std::max(a[1] * a[2], smax * smax * 1e-6f);
  }
  h = std::sqrt(s2 * occupancy / float(n));
  if (leaf_hint > 0.0f) {
    // One point per voxel of edge leaf_hint on surfaces: a cell of edge h
    // holds about (h / leaf)^2 of them.
    h = leaf_hint * std::sqrt(occupancy);
  }
  return true;
}

// Grows h until the dense grid fits the cell cap.
static float edge_fit_cap(const uint32_t* hb, float h, uint32_t n) {
  for (int guard = 0; guard < 64; ++guard) {
    GridDesc d;
    grid_desc(hb, h, n, d);
    if (uint64_t(d.dx) * uint64_t(d.dy) * uint64_t(d.dz) <= kMaxCells) break;
    h *= 1.25f;
  }
  return h;
}

// ---- edge choice from distinct-cell sketches --------------------------------
//
// The occupancy of non-empty cells at edge h is n / (number of distinct cells),
// which for a sparse cloud does not follow the h^2 surface law near the point
// spacing, so a guess has to be measured. One pass over the points measures it
// at kSketchEdges candidate edges at once: every point files itself into the
// cell it falls in at each edge and updates a HyperLogLog sketch (Flajolet et
// al., 2007) of that edge's distinct cells in LDS. A workgroup's registers are
// stored as bytes; one workgroup per cloud takes the register-wise maximum over
// its workgroups (order-free, so deterministic) and turns it into estimates.
// The register arrays are small, so no atomics go to device memory (counting
// distinct cells exactly needs a cell-sized array of device-wide atomics).

constexpr int kSketchEdges = 8;         // candidate edges per cloud
constexpr int kSketchRegs = 256;        // registers per sketch: ~6.5% standard error
constexpr uint32_t kSketchChunk = 4096; // points per workgroup, at least
// Workgroups per cloud, at most: enough that the two clouds of one
// registration fill the chip's 256 CUs (at 64 a 0.9M-point cloud took 72 us).
// The merged registers, and so the estimates, do not depend on the split.
constexpr uint32_t kSketchMaxWg = 256;
constexpr float kSketchStep = 1.6f;     // ratio of consecutive candidate edges
constexpr int kSketchBelow = 2;         // candidates below the first guess

struct SketchSeg {
  float ox, oy, oz;
  float inv_h[kSketchEdges];
  uint32_t pt_off;  // first point
  uint32_t n;       // points
  uint32_t wg_off;  // first workgroup
  uint32_t nwg;     // workgroups
  uint32_t chunk;   // points per workgroup
};

// 64-bit finalizer of MurmurHash3 (a bijection that spreads every input bit).
__device__ __forceinline__ unsigned long long sketch_mix(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

__device__ __forceinline__ unsigned long long sketch_coord(float v, float o, float inv_h) {
  const float f = floorf((v - o) * inv_h);
  return f <= 0.0f ? 0ull : (f >= 2097151.0f ? 2097151ull : (unsigned long long)f);
}

__global__ void __launch_bounds__(kBlock) k_edge_sketch(const float4* __restrict__ pts,
                                                        const SketchSeg* __restrict__ segs,
                                                        uint32_t nseg,
                                                        const uint32_t* __restrict__ wg_begin,
                                                        uint8_t* __restrict__ regs_out) {
  __shared__ uint32_t reg[kSketchEdges][kSketchRegs];
  for (int t = threadIdx.x; t < kSketchEdges * kSketchRegs; t += blockDim.x) (&reg[0][0])[t] = 0u;
  __syncthreads();
  const uint32_t s = seg_of(wg_begin, nseg, blockIdx.x);
  const SketchSeg& S = segs[s];
  const uint32_t first = (blockIdx.x - S.wg_off) * S.chunk;
  const uint32_t last = min(first + S.chunk, S.n);
  for (uint32_t i = first + threadIdx.x; i < last; i += blockDim.x) {
    const float4 p = pts[S.pt_off + i];
#pragma unroll
    for (int j = 0; j < kSketchEdges; ++j) {
      const float ih = S.inv_h[j];
      const unsigned long long key = sketch_coord(p.x, S.ox, ih) |
                                     (sketch_coord(p.y, S.oy, ih) << 21) |
                                     (sketch_coord(p.z, S.oz, ih) << 42);
      const unsigned long long hv = sketch_mix(key);
      const uint32_t idx = uint32_t(hv >> 56);
      const unsigned long long rest = hv << 8;
      const uint32_t rho = rest ? uint32_t(__clzll((long long)rest)) + 1u : 57u;
      atomicMax(&reg[j][idx], rho);
    }
  }
  __syncthreads();
  uint8_t* out = regs_out + size_t(blockIdx.x) * (kSketchEdges * kSketchRegs);
  for (int t = threadIdx.x; t < kSketchEdges * kSketchRegs; t += blockDim.x)
    out[t] = uint8_t((&reg[0][0])[t]);
}

// One workgroup per (cloud, candidate edge), a thread per register: the
// estimate of distinct cells, with the small-range correction (linear counting).
__global__ void __launch_bounds__(kSketchRegs) k_edge_sketch_merge(
    const uint8_t* __restrict__ regs, const SketchSeg* __restrict__ segs,
    float* __restrict__ est) {
  const uint32_t seg = blockIdx.x / kSketchEdges, j = blockIdx.x % kSketchEdges;
  const SketchSeg& S = segs[seg];
  __shared__ double inv_sum[kSketchRegs];
  __shared__ uint32_t zeros[kSketchRegs];
  const uint32_t m = threadIdx.x;
  const uint8_t* base = regs + (size_t(S.wg_off) * kSketchEdges + j) * kSketchRegs + m;
  uint32_t r = 0;
#pragma unroll 8
  for (uint32_t w = 0; w < S.nwg; ++w)
    r = max(r, uint32_t(base[size_t(w) * (kSketchEdges * kSketchRegs)]));
  inv_sum[m] = ldexp(1.0, -int(r));
  zeros[m] = r == 0 ? 1u : 0u;
  __syncthreads();
  for (uint32_t h = kSketchRegs / 2; h > 0; h >>= 1) {
    if (m < h) {
      inv_sum[m] += inv_sum[m + h];
      zeros[m] += zeros[m + h];
    }
    __syncthreads();
  }
  if (m == 0) {
    const double M = double(kSketchRegs);
    const double alpha = 0.7213 / (1.0 + 1.079 / M);
    double e = alpha * M * M / inv_sum[0];
    if (e <= 2.5 * M && zeros[0] > 0) e = M * log(M / double(zeros[0]));
    est[blockIdx.x] = float(e);
  }
}

// Candidate edges around a first guess g: g * kSketchStep^(j - kSketchBelow).
static void sketch_candidates(float g, float* hc) {
  for (int j = 0; j < kSketchEdges; ++j)
    hc[j] = g * std::pow(kSketchStep, float(j - kSketchBelow));
}

// The edge whose estimated occupancy n / cells meets `occupancy`: log-log
// interpolation between the bracketing candidates; beyond the candidates, the
// surface law (occupancy ~ h^2) from the nearest one, within a factor of 4.
static float edge_from_sketch(uint32_t n, const float* hc, const float* est, float occupancy) {
  float o[kSketchEdges];
  for (int j = 0; j < kSketchEdges; ++j) o[j] = float(n) / std::max(est[j], 1.0f);
  if (o[0] >= occupancy) {
    const float f = std::max(std::sqrt(occupancy / o[0]), 0.25f);
    return hc[0] * f;
  }
  for (int j = 0; j + 1 < kSketchEdges; ++j) {
    if (o[j + 1] < occupancy) continue;
    const float lo = std::log(std::max(o[j], 1e-6f)), hi = std::log(o[j + 1]);
    if (!(hi > lo)) return hc[j + 1];
    const float t = (std::log(occupancy) - lo) / (hi - lo);
    return hc[j] * std::pow(kSketchStep, std::min(std::max(t, 0.0f), 1.0f));
  }
  const float f = std::min(std::sqrt(occupancy / std::max(o[kSketchEdges - 1], 1e-6f)), 4.0f);
  return hc[kSketchEdges - 1] * f;
}

// Sketches segments `ids` of a cloud at their candidate edges (hc: kSketchEdges
// per listed segment) and returns the chosen edges in h[ids[k]].
static void sketch_edges(const float4* pts, const std::vector<uint32_t>& off,
                         const std::vector<uint32_t>& hb, const std::vector<uint32_t>& ids,
                         const std::vector<float>& hc, float occupancy, std::vector<float>& h,
                         hipStream_t s) {
  const uint32_t m = uint32_t(ids.size());
  if (m == 0) return;
  std::vector<SketchSeg> segs(m);
  std::vector<uint32_t> wg_begin(m);
  uint32_t wgs = 0;
  for (uint32_t k = 0; k < m; ++k) {
    const uint32_t g = ids[k];
    SketchSeg& S = segs[k];
    S.ox = ord2f(hb[6 * g + 0]);
    S.oy = ord2f(hb[6 * g + 1]);
    S.oz = ord2f(hb[6 * g + 2]);
    for (int j = 0; j < kSketchEdges; ++j) S.inv_h[j] = 1.0f / hc[size_t(k) * kSketchEdges + j];
    S.pt_off = off[g];
    S.n = off[g + 1] - off[g];
    S.wg_off = wgs;
    S.nwg = std::min(std::max<uint32_t>(1, (S.n + kSketchChunk - 1) / kSketchChunk), kSketchMaxWg);
    S.chunk = (S.n + S.nwg - 1) / S.nwg;
    wg_begin[k] = wgs;
    wgs += S.nwg;
  }
  HostPack pk;
  const size_t o_segs = pk.add(segs), o_wg = pk.add(wg_begin);
  DevBuf<uint8_t> tables(pk.bytes.size());
  DevBuf<uint8_t> regs(size_t(wgs) * kSketchEdges * kSketchRegs);
  DevBuf<float> est(size_t(m) * kSketchEdges);
  upload(tables.p, pk.bytes.data(), pk.bytes.size(), s);  // one copy for both tables
  const SketchSeg* d_segs = HostPack::at<SketchSeg>(tables.p, o_segs);
  hipLaunchKernelGGL(k_edge_sketch, dim3(wgs), dim3(kBlock), 0, s, pts, d_segs, m,
                     HostPack::at<uint32_t>(tables.p, o_wg), regs.p);
  hipLaunchKernelGGL(k_edge_sketch_merge, dim3(m * kSketchEdges), dim3(kSketchRegs), 0, s, regs.p,
                     d_segs, est.p);
  std::vector<float> host(size_t(m) * kSketchEdges);
  download(host.data(), est.p, sizeof(float) * host.size(), s);  // also keeps the tables alive until the kernels end
  for (uint32_t k = 0; k < m; ++k)
    h[ids[k]] = edge_from_sketch(segs[k].n, &hc[size_t(k) * kSketchEdges],
                                 &host[size_t(k) * kSketchEdges], occupancy);
}

float choose_cell_edge(const Cloud& c, const uint32_t* hb, float occupancy, float leaf_hint) {
  std::vector<float> h(1, 1.0f);
  if (!edge_first_guess(hb, c.n, occupancy, leaf_hint, h[0])) return h[0];
  std::vector<float> hc(kSketchEdges);
  sketch_candidates(h[0], hc.data());
  const std::vector<uint32_t> off = {0u, c.n}, ids = {0u};
  const std::vector<uint32_t> hbv(hb, hb + 6);
  sketch_edges(c.pts.p, off, hbv, ids, hc, occupancy, h, thread_stream());
  return edge_fit_cap(hb, h[0], c.n);
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
  DevBuf<uint32_t> counts(size_t(nc) + 1), cid(n > 0 ? n : 1), rnk(n > 0 ? n : 1), scal(2);
  fill_u32(counts.p, 0u, size_t(nc) + 1, s);
  if (n > 0) {
    hipLaunchKernelGGL(k_grid_count, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, c.pts.p, n,
                       g.desc, counts.p, cid.p, rnk.p);
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
    hipLaunchKernelGGL(k_grid_rowbits, dim3(uint32_t((32 * nrw + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       g.occ.p, g.desc, uint32_t(nrw), g.rowocc.p);
  }
  if (n == 0) return;
  hipLaunchKernelGGL(k_grid_scatter, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, cid.p,
                     rnk.p, n, g.cell_start.p, g.order.p, scal.p);  // (scal: big and medium cell counts)
  DevBuf<uint32_t> big(std::max<uint32_t>(nc, 1));
  hipLaunchKernelGGL(k_sort_cells, dim3((nc + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     g.cell_start.p, nc, g.order.p, big.p, scal.p);
  hipLaunchKernelGGL(k_sort_medium_cells, dim3(512), dim3(kBlock), 0, s, g.cell_start.p, nc,
                     g.order.p, big.p, scal.p);
  hipLaunchKernelGGL(k_sort_big_cells, dim3(128), dim3(kBlock), 0, s, g.cell_start.p, g.order.p,
                     big.p, scal.p);
  hipLaunchKernelGGL(k_grid_gather, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, c.pts.p,
                     g.order.p, n, g.pts.p);
  // counts/cid/rnk/big return to the cache; ordering on s keeps reuse safe.
}

// Chebyshev distance transform of occupancy, separable: along one axis
// d'(c) = min_k max(|k|, d(c + k e_axis)) for |k| <= cap; three passes give the
// min over the cube. Values saturate at cap + 1 ("at least cap + 1").
__global__ void k_empty_init(const uint32_t* __restrict__ cs, uint32_t ncells,
                             uint8_t* __restrict__ out) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) out[c] = cs[c + 1] > cs[c] ? 0 : uint8_t(kEmptyCap + 1);
}

__device__ __forceinline__ void empty_pass_cell(const uint8_t* __restrict__ in,
                                                uint8_t* __restrict__ out, const GridDesc& g,
                                                int axis, uint32_t c) {
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

__global__ void k_empty_pass(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, GridDesc g,
                             int axis) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < g.ncells) empty_pass_cell(in, out, g, axis, c);
}

void chebyshev_field(uint8_t* field, const GridDesc& d, hipStream_t s) {
  const uint32_t nc = d.ncells;
  if (nc == 0) return;
  DevBuf<uint8_t> tmp(nc);
  const dim3 grid((nc + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_empty_pass, grid, dim3(kBlock), 0, s, field, tmp.p, d, 0);
  hipLaunchKernelGGL(k_empty_pass, grid, dim3(kBlock), 0, s, tmp.p, field, d, 1);
  hipLaunchKernelGGL(k_empty_pass, grid, dim3(kBlock), 0, s, field, tmp.p, d, 2);
  SR_HIP(hipMemcpyAsync(field, tmp.p, nc, hipMemcpyDeviceToDevice, s));
}

void build_empty_radius(Grid& g, hipStream_t s) {
  const uint32_t nc = g.desc.ncells;
  g.empty_radius.resize(std::max<uint32_t>(nc, 1));
  if (nc == 0) return;
  const dim3 grid((nc + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_empty_init, grid, dim3(kBlock), 0, s, g.cell_start.p, nc, g.empty_radius.p);
  chebyshev_field(g.empty_radius.p, g.desc, s);
}

// ---- many clouds in one set of launches -----------------------------------
//
// Segment s is points [off[s], off[s+1]) of one cloud; each gets its own grid
// (edge, origin, dims) and the grids share flat arrays: a segment's cells are
// ids [cell_off[s], cell_off[s+1]) of one counting sort, whose positions equal
// the points' (so a segment's points stay in [off[s], off[s+1]) in grid order),
// and its cell_start (ncells + 1 entries at cs_off[s]) counts from 0 within
// the segment, as the one-cloud build does.

struct SegGrid {
  GridDesc d;         // bits/rows already point at the segment's words
  uint32_t pt_off;    // first point
  uint32_t n;         // points
  uint32_t cell_off;  // first cell id
  uint32_t cs_off;    // first cell_start entry
  uint32_t word_off;  // first occupancy word
  uint32_t row_off;   // first row-bitmap word
  uint32_t pad_[2];
};

__global__ void k_seg_grid_count(const float4* __restrict__ pts, const SegGrid* __restrict__ sg,
                                 uint32_t nseg, const uint32_t* __restrict__ blk_begin,
                                 uint32_t* __restrict__ counts, uint32_t* __restrict__ cid,
                                 uint32_t* __restrict__ rnk) {
  const uint32_t s = seg_of(blk_begin, nseg, blockIdx.x);
  const SegGrid& G = sg[s];
  const uint32_t il = (blockIdx.x - blk_begin[s]) * blockDim.x + threadIdx.x;
  if (il >= G.n) return;
  const GridDesc& g = G.d;
  const float4 p = pts[G.pt_off + il];
  int cx = cell_coord(p.x, g.ox, g.inv_h, g.dx);
  int cy = cell_coord(p.y, g.oy, g.inv_h, g.dy);
  int cz = cell_coord(p.z, g.oz, g.inv_h, g.dz);
  uint32_t c = G.cell_off + (uint32_t(cz) * uint32_t(g.dy) + uint32_t(cy)) * uint32_t(g.dx) + uint32_t(cx);
  rnk[G.pt_off + il] = atomicAdd(&counts[c], 1u);
  cid[G.pt_off + il] = c;
}

// Per segment: cell_start entries from the flat exclusive scan, minus the
// segment's first point.
__global__ void k_seg_local_cs(const uint32_t* __restrict__ csg, const SegGrid* __restrict__ sg,
                               uint32_t nseg, const uint32_t* __restrict__ cs_begin, uint32_t total,
                               uint32_t* __restrict__ cs_local) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint32_t s = seg_of(cs_begin, nseg, t);
  const SegGrid& G = sg[s];
  cs_local[t] = csg[G.cell_off + (t - G.cs_off)] - G.pt_off;
}

__global__ void k_seg_grid_bits(const uint32_t* __restrict__ cs_local, const SegGrid* __restrict__ sg,
                                uint32_t nseg, const uint32_t* __restrict__ word_begin,
                                uint32_t total, uint32_t* __restrict__ occ) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint32_t s = seg_of(word_begin, nseg, t);
  const SegGrid& G = sg[s];
  grid_bits_word(cs_local + G.cs_off, G.d, t - G.word_off, occ + G.word_off);
}

// Row bitmaps of every segment, one thread per row as in k_grid_rowbits
// (thread t serves bit t % 32 of word t / 32 of the flat row-bitmap array).
__global__ void k_seg_grid_rowbits(const uint32_t* __restrict__ occ, const SegGrid* __restrict__ sg,
                                   uint32_t nseg, const uint32_t* __restrict__ row_begin,
                                   uint32_t total, uint32_t* __restrict__ rows) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t wi = t >> 5;
  uint32_t any = 0;
  if (wi < total) {
    const SegGrid& G = sg[seg_of(row_begin, nseg, wi)];
    const GridDesc& g = G.d;
    const uint32_t w = wi - G.row_off;
    const uint32_t cz = w / g.wps, cy = (w % g.wps) * 32u + (t & 31u);
    if (cy < uint32_t(g.dy)) {
      const uint32_t* rb = occ + G.word_off + (size_t(cz) * uint32_t(g.dy) + cy) * g.wpr;
      for (uint32_t k = 0; k < g.wpr; ++k) any |= rb[k];
    }
  }
  const unsigned long long m = __ballot(any != 0u);
  const uint32_t lane = threadIdx.x & 63u;
  if ((lane & 31u) == 0u && wi < total) rows[wi] = uint32_t(m >> lane);
}

__global__ void k_seg_grid_gather(const float4* __restrict__ pts, const uint32_t* __restrict__ order,
                                  const uint32_t* __restrict__ pt_begin, uint32_t nseg, uint32_t n,
                                  float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = seg_of(pt_begin, nseg, i);
  const uint32_t o = order[i];
  const float4 p = pts[o];
  out[i] = make_float4(p.x, p.y, p.z, __uint_as_float(o - pt_begin[s]));
}

__global__ void k_seg_empty_init(const uint32_t* __restrict__ cs_local, const SegGrid* __restrict__ sg,
                                 uint32_t nseg, const uint32_t* __restrict__ cell_begin,
                                 uint32_t total, uint8_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint32_t s = seg_of(cell_begin, nseg, t);
  const SegGrid& G = sg[s];
  const uint32_t* cs = cs_local + G.cs_off;
  const uint32_t c = t - G.cell_off;
  out[t] = cs[c + 1] > cs[c] ? 0 : uint8_t(kEmptyCap + 1);
}

__global__ void k_seg_empty_pass(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                 const SegGrid* __restrict__ sg, uint32_t nseg,
                                 const uint32_t* __restrict__ cell_begin, uint32_t total, int axis) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint32_t s = seg_of(cell_begin, nseg, t);
  const SegGrid& G = sg[s];
  empty_pass_cell(in + G.cell_off, out + G.cell_off, G.d, axis, t - G.cell_off);
}

static std::vector<uint32_t> blocks_of(const std::vector<uint32_t>& off) {
  std::vector<uint32_t> b(off.size());
  b[0] = 0;
  for (size_t s = 0; s + 1 < off.size(); ++s)
    b[s + 1] = b[s] + std::max<uint32_t>(1, (off[s + 1] - off[s] + kBlock - 1) / kBlock);
  return b;
}

void choose_cell_edges_many(const Cloud& c, const std::vector<uint32_t>& off,
                            const std::vector<uint32_t>& hb, float occupancy, float leaf_hint,
                            std::vector<float>& h, hipStream_t s) {
  // choose_cell_edge per segment, with the sketches of all segments in one launch
  const uint32_t nseg = uint32_t(off.size() - 1);
  h.assign(nseg, 1.0f);
  std::vector<uint32_t> ids;
  std::vector<float> hc;
  for (uint32_t g = 0; g < nseg; ++g) {
    if (!edge_first_guess(&hb[6 * g], off[g + 1] - off[g], occupancy, leaf_hint, h[g])) continue;
    ids.push_back(g);
    hc.resize(hc.size() + kSketchEdges);
    sketch_candidates(h[g], &hc[hc.size() - kSketchEdges]);
  }
  sketch_edges(c.pts.p, off, hb, ids, hc, occupancy, h, s);
  for (uint32_t k = 0; k < ids.size(); ++k) {
    const uint32_t g = ids[k];
    h[g] = edge_fit_cap(&hb[6 * g], h[g], off[g + 1] - off[g]);
  }
}

void build_grids_many(const Cloud& c, const std::vector<uint32_t>& off,
                      const std::vector<uint32_t>& hb, const std::vector<float>& h,
                      uint32_t empty_segs, GridSet& out, hipStream_t s) {
  const uint32_t nseg = uint32_t(off.size() - 1), n = off.back();
  out.desc.assign(nseg, GridDesc{});
  out.pt_off = off;
  out.cell_off.assign(nseg + 1, 0);
  out.cs_off.assign(nseg, 0);
  std::vector<SegGrid> sg(nseg);
  std::vector<uint32_t> word_begin(nseg), row_begin(nseg), cs_begin(nseg);
  uint64_t cells = 0, words = 0, rws = 0;
  for (uint32_t g = 0; g < nseg; ++g) {
    SegGrid& G = sg[g];
    grid_desc(&hb[6 * g], h[g], off[g + 1] - off[g], G.d);
    const uint64_t nc = uint64_t(G.d.dx) * uint64_t(G.d.dy) * uint64_t(G.d.dz);
    if (nc > kMaxCells)
      fail(3, "grid would need " + std::to_string(nc) + " cells (cap " + std::to_string(kMaxCells) +
                  "); use a larger radius or the CPU path");
    G.pt_off = off[g];
    G.n = off[g + 1] - off[g];
    G.cell_off = uint32_t(cells);
    G.cs_off = uint32_t(cells + g);
    G.word_off = uint32_t(words);
    G.row_off = uint32_t(rws);
    out.cell_off[g] = G.cell_off;
    out.cs_off[g] = G.cs_off;
    word_begin[g] = G.word_off;
    row_begin[g] = G.row_off;
    cs_begin[g] = G.cs_off;
    cells += G.d.ncells;
    words += uint64_t(G.d.wpr) * uint64_t(G.d.dy) * uint64_t(G.d.dz);
    rws += uint64_t(G.d.wps) * uint64_t(G.d.dz);
    if (cells + nseg > 0xFFFFFFFFull || words > 0xFFFFFFFFull)
      fail(3, "too many grid cells for one batch");
  }
  out.cell_off[nseg] = uint32_t(cells);
  const uint32_t ncells = uint32_t(cells), ncs = uint32_t(cells + nseg);
  out.cell_start.resize(std::max<uint32_t>(ncs, 1));
  out.pts.resize(std::max<uint32_t>(n, 1));
  out.occ.resize(std::max<uint64_t>(words, 1));
  out.rowocc.resize(std::max<uint64_t>(rws, 1));
  for (uint32_t g = 0; g < nseg; ++g) {
    sg[g].d.bits = out.occ.p + sg[g].word_off;
    sg[g].d.rows = out.rowocc.p + sg[g].row_off;
    out.desc[g] = sg[g].d;
  }
  std::vector<uint32_t> blk = blocks_of(off);
  std::vector<uint32_t> cell_begin(out.cell_off.begin(), out.cell_off.end() - 1);
  // the per-segment tables go up in one copy (each small copy costs a few
  // microseconds of its own on the stream)
  HostPack pk;
  const size_t o_sg = pk.add(sg), o_blk = pk.add(blk), o_word = pk.add(word_begin),
               o_row = pk.add(row_begin), o_cs = pk.add(cs_begin), o_cell = pk.add(cell_begin),
               o_pt = pk.add(off);
  DevBuf<uint8_t> tables(pk.bytes.size());
  upload(tables.p, pk.bytes.data(), pk.bytes.size(), s);
  const SegGrid* d_sg = HostPack::at<SegGrid>(tables.p, o_sg);
  const uint32_t *d_blk = HostPack::at<uint32_t>(tables.p, o_blk),
                 *d_word = HostPack::at<uint32_t>(tables.p, o_word),
                 *d_row = HostPack::at<uint32_t>(tables.p, o_row),
                 *d_cs = HostPack::at<uint32_t>(tables.p, o_cs),
                 *d_cell = HostPack::at<uint32_t>(tables.p, o_cell),
                 *d_pt = HostPack::at<uint32_t>(tables.p, o_pt);
  DevBuf<uint32_t> counts(size_t(ncells) + 1), csg(size_t(ncells) + 1), cid(std::max<uint32_t>(n, 1)),
      rnk(std::max<uint32_t>(n, 1)), order(std::max<uint32_t>(n, 1)), scal(2);
  // The scratch below is released at the stream's next wait instead of
  // waiting here (inside the allocation-free entry point, where that is not
  // available, the stream is waited for at the end).
  bool deferred = true;
  fill_u32(counts.p, 0u, size_t(ncells) + 1, s);
  hipLaunchKernelGGL(k_seg_grid_count, dim3(blk.back()), dim3(kBlock), 0, s, c.pts.p, d_sg, nseg,
                     d_blk, counts.p, cid.p, rnk.p);
  exclusive_scan_u32(counts.p, csg.p, ncells + 1, nullptr, s);
  auto grid1 = [](uint64_t t) { return dim3(uint32_t(std::max<uint64_t>((t + kBlock - 1) / kBlock, 1))); };
  hipLaunchKernelGGL(k_seg_local_cs, grid1(ncs), dim3(kBlock), 0, s, csg.p, d_sg, nseg, d_cs, ncs,
                     out.cell_start.p);
  if (words)
    hipLaunchKernelGGL(k_seg_grid_bits, grid1(words), dim3(kBlock), 0, s, out.cell_start.p, d_sg,
                       nseg, d_word, uint32_t(words), out.occ.p);
  if (rws)
    hipLaunchKernelGGL(k_seg_grid_rowbits, grid1(32 * rws), dim3(kBlock), 0, s, out.occ.p, d_sg, nseg,
                       d_row, uint32_t(rws), out.rowocc.p);
  if (n) {
    hipLaunchKernelGGL(k_grid_scatter, grid1(n), dim3(kBlock), 0, s, cid.p, rnk.p, n, csg.p, order.p,
                       scal.p);  // (scal: big and medium cell counts)
    DevBuf<uint32_t> big(std::max<uint32_t>(ncells, 1));
    hipLaunchKernelGGL(k_sort_cells, grid1(ncells), dim3(kBlock), 0, s, csg.p, ncells, order.p,
                       big.p, scal.p);
    hipLaunchKernelGGL(k_sort_medium_cells, dim3(512), dim3(kBlock), 0, s, csg.p, ncells, order.p,
                       big.p, scal.p);
    hipLaunchKernelGGL(k_sort_big_cells, dim3(128), dim3(kBlock), 0, s, csg.p, order.p, big.p, scal.p);
    hipLaunchKernelGGL(k_seg_grid_gather, grid1(n), dim3(kBlock), 0, s, c.pts.p, order.p, d_pt, nseg,
                       n, out.pts.p);
    deferred = big.release_after(s) && deferred;
  }
  empty_segs = std::min(empty_segs, nseg);
  const uint32_t ecells = empty_segs ? out.cell_off[empty_segs] : 0u;
  if (ecells) {
    // ping-pong between two buffers; the third pass lands in `tmp`, which
    // then becomes the result
    DevBuf<uint8_t> er(ecells), tmp(ecells);
    hipLaunchKernelGGL(k_seg_empty_init, grid1(ecells), dim3(kBlock), 0, s, out.cell_start.p, d_sg,
                       empty_segs, d_cell, ecells, er.p);
    for (int axis = 0; axis < 3; ++axis) {
      const uint8_t* in = axis == 1 ? tmp.p : er.p;
      uint8_t* o = axis == 1 ? er.p : tmp.p;
      hipLaunchKernelGGL(k_seg_empty_pass, grid1(ecells), dim3(kBlock), 0, s, in, o, d_sg,
                         empty_segs, d_cell, ecells, axis);
    }
    out.empty_radius = std::move(tmp);
    deferred = er.release_after(s) && deferred;
  }
  for (DevBuf<uint32_t>* b : {&counts, &csg, &cid, &rnk, &order, &scal})
    deferred = b->release_after(s) && deferred;
  deferred = tables.release_after(s) && deferred;
  if (!deferred) SR_HIP(hipStreamSynchronize(s));
}

}  // namespace sr
