// Shared device-side definitions for the gfx950 GICP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sr {

constexpr int kWave = 64;     // CDNA wavefront width
constexpr int kBlock = 256;   // threads per workgroup for point-parallel kernels
constexpr uint32_t kNone = 0xFFFFFFFFu;

// Optional search counters (profiling): summed over all queries of a launch.
// Work counters of a grid search (search-statistics mode). outcome[]: queries
// rejected without a scan, finished in the 3x3x3 block, finished in the ring
// search with a neighbour, finished in the ring search without one.
enum { kRejected = 0, kBlockDone = 1, kRingFound = 2, kRingNone = 3 };
struct SearchStats {
  unsigned long long queries, rings, cells, points;
  unsigned long long outcome[4];
  unsigned long long max_cells, max_points;  // largest single-query counts
};

__device__ __forceinline__ void stats_add(SearchStats* st, uint32_t rings, uint32_t cells,
                                          uint32_t points, int outcome) {
  if (!st) return;
  atomicAdd(&st->queries, 1ull);
  atomicAdd(&st->rings, (unsigned long long)rings);
  atomicAdd(&st->cells, (unsigned long long)cells);
  atomicAdd(&st->points, (unsigned long long)points);
  atomicAdd(&st->outcome[outcome], 1ull);
  atomicMax(&st->max_cells, (unsigned long long)cells);
  atomicMax(&st->max_points, (unsigned long long)points);
}

// Uniform grid description. Cell (cx,cy,cz) has linear id (cz*dy + cy)*dx + cx,
// the layout of build_grid (crates/spatialrust-search/src/uniform_grid.rs:56-64).
struct GridDesc {
  float ox, oy, oz;  // origin (per-axis minimum of the points)
  float h;           // cell edge
  float inv_h;       // 1/h in f32, as the reference computes inv_cell
  int32_t dx, dy, dz;
  uint32_t ncells;
  uint32_t npoints;
  float margin;  // slack covering f32 cell-assignment rounding (see cell_lb2)
  uint32_t wpr;  // 32-bit words per x-row of the occupancy bitmap
  const uint32_t* bits;  // occupancy bitmap: row (cz*dy+cy) has wpr words, bit cx
  const uint32_t* rows;  // row bitmap: slab cz has wps words, bit cy = row has a point
  uint32_t wps;          // 32-bit words per z-slab of the row bitmap
  uint32_t pad2_;
};

// Segment of x in a launch over several segments: the last s < nseg with
// begins[s] <= x (begins ascending, begins[0] = 0).
__device__ __forceinline__ uint32_t seg_of(const uint32_t* __restrict__ begins, uint32_t nseg,
                                           uint32_t x) {
  uint32_t lo = 0, hi = nseg;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (begins[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Bitonic sorting network over N (a power of two) values in registers; every
// index is a compile-time constant once unrolled.
template <int N>
__device__ __forceinline__ void sort_net(uint32_t (&v)[N]) {
#pragma unroll
  for (int k = 2; k <= N; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const uint32_t a = v[i], b = v[l];
          const bool up = (i & k) == 0;
          v[i] = up ? min(a, b) : max(a, b);
          v[l] = up ? max(a, b) : min(a, b);
        }
      }
}

// Order-preserving map of f32 onto u32 so atomicMin/atomicMax work on floats.
__host__ __device__ __forceinline__ uint32_t f2ord(float f) {
#ifdef __HIP_DEVICE_COMPILE__
  uint32_t u = __float_as_uint(f);
#else
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
#endif
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float ord2f(uint32_t u) {
  uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
#ifdef __HIP_DEVICE_COMPILE__
  return __uint_as_float(b);
#else
  float f;
  __builtin_memcpy(&f, &b, 4);
  return f;
#endif
}

// Cell coordinate with the clamping of build_grid (uniform_grid.rs:56-62):
// floor((v - o) * inv) clamped to [0, dim-1].
__device__ __forceinline__ int cell_coord(float v, float o, float inv, int dim) {
  float f = floorf(__fmul_rn(__fsub_rn(v, o), inv));
  int c;
  if (!(f == f)) c = 0;
  else if (f < -1.0e9f) c = -1000000000;
  else if (f > 1.0e9f) c = 1000000000;
  else c = (int)f;
  return c < 0 ? 0 : (c >= dim ? dim - 1 : c);
}

// Squared distance exactly as squared_distance (kdtree.rs:544-549): three
// separately rounded products summed left to right, no FMA.
__device__ __forceinline__ float dist2(float px, float py, float pz, float qx, float qy, float qz) {
  float dx = __fsub_rn(px, qx), dy = __fsub_rn(py, qy), dz = __fsub_rn(pz, qz);
  return __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
}

// Unclamped cell coordinate of an arbitrary query (may lie outside the grid).
__device__ __forceinline__ int cell_coord_free(float v, float o, float inv) {
  float f = floorf((v - o) * inv);
  if (!(f == f)) return 0;
  f = fminf(fmaxf(f, -1.0e8f), 1.0e8f);
  return (int)f;
}

// Lower bound of the squared distance from p to any point filed in cell
// (cx,cy,cz). The cell box is widened by g.margin on every side: a point's cell
// comes from an f32 floor((v-o)*inv_h), whose rounding can file it up to a small
// fraction of a cell outside the exact box, so the bound must not exceed the
// distance of such a point.
__device__ __forceinline__ float cell_lb2(const GridDesc& g, float px, float py, float pz, int cx,
                                          int cy, int cz) {
  float lo, d, acc;
  lo = g.ox + (float)cx * g.h - g.margin;
  d = fmaxf(fmaxf(lo - px, px - (lo + g.h + 2.0f * g.margin)), 0.0f);
  acc = d * d;
  lo = g.oy + (float)cy * g.h - g.margin;
  d = fmaxf(fmaxf(lo - py, py - (lo + g.h + 2.0f * g.margin)), 0.0f);
  acc += d * d;
  lo = g.oz + (float)cz * g.h - g.margin;
  d = fmaxf(fmaxf(lo - pz, pz - (lo + g.h + 2.0f * g.margin)), 0.0f);
  acc += d * d;
  return acc;
}

// Distance from p to the outside of the block of cells within Chebyshev ring r
// of cell (hx,hy,hz), minus the filing margin: every point filed in a cell
// beyond ring r is at least this far from p.
__device__ __forceinline__ float ring_clearance(const GridDesc& g, float px, float py, float pz,
                                               int hx, int hy, int hz, int r) {
  float lo = g.ox + (float)(hx - r) * g.h, hi = g.ox + (float)(hx + r + 1) * g.h;
  float m = fminf(px - lo, hi - px);
  lo = g.oy + (float)(hy - r) * g.h;
  hi = g.oy + (float)(hy + r + 1) * g.h;
  m = fminf(m, fminf(py - lo, hi - py));
  lo = g.oz + (float)(hz - r) * g.h;
  hi = g.oz + (float)(hz + r + 1) * g.h;
  m = fminf(m, fminf(pz - lo, hi - pz));
  return m - g.margin;
}

// Distance from v to the widened interval of cell c along one axis (0 inside).
__device__ __forceinline__ float axis_gap(float v, float o, float h, float m, int c) {
  const float lo = o + (float)c * h - m;
  return fmaxf(fmaxf(lo - v, v - (lo + h + 2.0f * m)), 0.0f);
}

// Loads through the global address space: pointers read from descriptors in
// memory are generic, and generic (flat) loads also count against the LDS
// counter, so every wait on them waits for LDS traffic too.
// (float4 is read as a native vector: binding its class copy constructor
// would turn the pointer generic again.)
__device__ __forceinline__ uint32_t ldg(const uint32_t* p) {
  return *(const __attribute__((address_space(1))) uint32_t*)p;
}
__device__ __forceinline__ uint8_t ldg(const uint8_t* p) {
  return *(const __attribute__((address_space(1))) uint8_t*)p;
}
__device__ __forceinline__ float ldg(const float* p) {
  return *(const __attribute__((address_space(1))) float*)p;
}
// plain structs (no user-defined copy constructor)
template <typename T>
__device__ __forceinline__ T ldg(const T* p) {
#ifdef __HIP_DEVICE_COMPILE__
  T v;
  __builtin_memcpy(&v, (const __attribute__((address_space(1))) T*)p, sizeof(T));
  return v;
#else
  return *p;
#endif
}
__device__ __forceinline__ float4 ldg(const float4* p) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  const v4f v = *(const __attribute__((address_space(1))) v4f*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}

// Visits the points of N contiguous ranges [lo[k], hi[k]) of a grid-ordered
// array. W independent loads per step (indices clamped into the range, the
// extra lanes of a step skipped) keep several requests in flight per lane;
// W = 1 visits one point per step, for visitors too large to inline W times.
template <int W = 4, int N, typename VisitPoint>
__device__ __forceinline__ void stream_ranges(const uint32_t (&lo)[N], const uint32_t (&hi)[N],
                                              const float4* __restrict__ pts, VisitPoint&& visit) {
  static_assert(W >= 1 && W <= 16 && (W & (W - 1)) == 0, "stream width");
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (W == 1) {
      for (uint32_t s = lo[k]; s < hi[k]; ++s) visit(s, ldg(pts + s));
      continue;
    }
    for (uint32_t s = lo[k]; s < hi[k]; s += W) {
      const uint32_t last = hi[k] - 1;
      float4 q[W];
#pragma unroll
      for (int w = 0; w < W; ++w) q[w] = ldg(pts + min(s + uint32_t(w), last));
      visit(s, q[0]);
#pragma unroll
      for (int w = 1; w < W; ++w)
        if (s + uint32_t(w) <= last) visit(s + uint32_t(w), q[w]);
    }
  }
}

// Visits the points of N contiguous ranges [lo[k], hi[k]) as one sequence:
// each step loads W consecutive entries of the concatenation, so short ranges
// share the steps and the loads of several ranges are in flight together.
// Each range is padded to a multiple of W in the concatenation, so a step lies
// in one range: its range is picked once per step (a compare chain on the
// padded starts), not once per entry, and the padding entries are skipped.
template <int W, int N, typename VisitPoint>
__device__ __forceinline__ void stream_concat(const uint32_t (&lo)[N], const uint32_t (&hi)[N],
                                              const float4* __restrict__ pts, VisitPoint&& visit) {
  uint32_t start[N], delta[N], end[N];
  uint32_t total = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const uint32_t len = hi[k] - lo[k];
    start[k] = total;
    delta[k] = lo[k] - total;  // position = offset + delta (mod 2^32)
    end[k] = total + len;      // first padding offset of range k
    total += (len + uint32_t(W - 1)) & ~uint32_t(W - 1);
  }
  for (uint32_t c = 0; c < total; c += W) {
    // the last range starting at or before c (an empty range's successor
    // starts at the same offset and wins)
    uint32_t d = delta[0], e = end[0];
#pragma unroll
    for (int k = 1; k < N; ++k) {
      const bool in = c >= start[k];
      d = in ? delta[k] : d;
      e = in ? end[k] : e;
    }
    uint32_t pos[W];
    float4 q[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      pos[w] = min(c + uint32_t(w), e - 1u) + d;  // padding reloads the range's last point
      q[w] = ldg(pts + pos[w]);
    }
#pragma unroll
    for (int w = 0; w < W; ++w)
      if (c + uint32_t(w) < e) visit(pos[w], q[w]);
  }
}

// Visits every point filed in Chebyshev ring r around (hx,hy,hz) whose cell can
// hold a point within sqrt(thr) of p (thr is re-read as `visit` tightens it).
// Slabs and rows outside the sphere are cut analytically and empty rows are
// skipped through the slab's row bitmap. In grid order a row's cells are
// adjacent, so a face row of the ring is one contiguous point range (its x span
// cut to the sphere) and an interior row contributes its two end cells; each
// row's range bounds are independent loads and its points are streamed.
// `rows` counts the rows scanned; W is the streaming width (stream_ranges).
template <int W = 4, typename VisitPoint>
__device__ __forceinline__ void scan_ring(const GridDesc& g, const uint32_t* __restrict__ cs,
                                          const float4* __restrict__ pts, float px, float py,
                                          float pz, int hx, int hy, int hz, int r, const float& thr,
                                          uint32_t& rows, VisitPoint&& visit) {
  const int z0 = max(hz - r, 0), z1 = min(hz + r, g.dz - 1);
  for (int cz = z0; cz <= z1; ++cz) {
    const float gz = axis_gap(pz, g.oz, g.h, g.margin, cz);
    const float gz2 = gz * gz;
    if (gz2 > thr) continue;
    const bool zedge = (cz == hz - r) || (cz == hz + r);
    int y0 = max(hy - r, 0), y1 = min(hy + r, g.dy - 1);
    if (thr < INFINITY) {
      // rows whose y-gap fits in the remaining radius (+1 cell of slack covers
      // the approximate sqrt and the filing margin)
      const float ry = __builtin_amdgcn_sqrtf(thr - gz2) + g.margin;
      y0 = max(y0, cell_coord_free(py - ry, g.oy, g.inv_h) - 1);
      y1 = min(y1, cell_coord_free(py + ry, g.oy, g.inv_h) + 1);
    }
    if (y0 > y1) continue;
    const uint32_t* sb = g.rows + size_t(cz) * g.wps;
    for (int wy = y0 >> 5; wy <= (y1 >> 5); ++wy) {
      const int ylo = max(y0 - wy * 32, 0), yhi = min(y1 - wy * 32, 31);
      uint32_t rm = sb[wy] & (0xFFFFFFFFu >> (31 - yhi)) & (0xFFFFFFFFu << ylo);
      while (rm) {  // occupied rows only
        const int cy = wy * 32 + (__ffs(rm) - 1);
        rm &= rm - 1;
        const float gy = axis_gap(py, g.oy, g.h, g.margin, cy);
        const float gyz = gz2 + gy * gy;
        if (gyz > thr) continue;
        const uint32_t cbase = (uint32_t(cz) * uint32_t(g.dy) + uint32_t(cy)) * uint32_t(g.dx);
        uint32_t lo[2] = {0, 0}, hi[2] = {0, 0};
        if (zedge || cy == hy - r || cy == hy + r) {
          int x0 = max(hx - r, 0), x1 = min(hx + r, g.dx - 1);
          if (thr < INFINITY) {
            const float rx = __builtin_amdgcn_sqrtf(thr - gyz) + g.margin;
            x0 = max(x0, cell_coord_free(px - rx, g.ox, g.inv_h) - 1);
            x1 = min(x1, cell_coord_free(px + rx, g.ox, g.inv_h) + 1);
          }
          if (x0 > x1) continue;
          lo[0] = ldg(cs + (cbase + uint32_t(x0)));
          hi[0] = ldg(cs + (cbase + uint32_t(x1) + 1));
        } else {
          // interior row of the ring: only its two end cells are new
          const int ends[2] = {hx - r, hx + r};
#pragma unroll
          for (int e = 0; e < 2; ++e) {
            const int cx = ends[e];
            if (cx < 0 || cx >= g.dx) continue;
            const float gx = axis_gap(px, g.ox, g.h, g.margin, cx);
            if (gyz + gx * gx > thr) continue;
            lo[e] = ldg(cs + (cbase + uint32_t(cx)));
            hi[e] = ldg(cs + (cbase + uint32_t(cx) + 1));
          }
        }
        ++rows;
        stream_ranges<W>(lo, hi, pts, visit);
      }
    }
  }
}

// Distance from (py,pz) to the outside of the (y,z) box of rows within
// Chebyshev distance r of row (hy,hz), minus the filing margin: a lower bound
// on the (y,z) distance to any point filed in a row farther than r.
__device__ __forceinline__ float clearance_yz(const GridDesc& g, float py, float pz, int hy, int hz,
                                              int r) {
  const float ylo = g.oy + (float)(hy - r) * g.h, yhi = g.oy + (float)(hy + r + 1) * g.h;
  const float zlo = g.oz + (float)(hz - r) * g.h, zhi = g.oz + (float)(hz + r + 1) * g.h;
  return fminf(fminf(py - ylo, yhi - py), fminf(pz - zlo, zhi - pz)) - g.margin;
}

// Visits every point filed in the 3x3x3 block of cells around (hx,hy,hz)
// (rings 0 and 1). In grid order the three cells of a block row are adjacent,
// so the block is nine contiguous point ranges; all eighteen range bounds are
// loaded up front (independent loads) and the points are then streamed, which
// keeps the chain of dependent memory round trips at two.
template <int W = 4, bool kConcat = false, typename VisitPoint>
__device__ __forceinline__ void scan_block27(const GridDesc& g, const uint32_t* __restrict__ cs,
                                             const float4* __restrict__ pts, int hx, int hy, int hz,
                                             VisitPoint&& visit) {
  const int x0 = max(hx - 1, 0), x1 = min(hx + 1, g.dx - 1);
  if (x0 > x1) return;
  uint32_t lo[9], hi[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int cz = hz - 1 + k / 3, cy = hy - 1 + k % 3;
    lo[k] = hi[k] = 0;
    if (cz >= 0 && cz < g.dz && cy >= 0 && cy < g.dy) {
      const uint32_t base = (uint32_t(cz) * uint32_t(g.dy) + uint32_t(cy)) * uint32_t(g.dx);
      lo[k] = ldg(cs + (base + uint32_t(x0)));
      hi[k] = ldg(cs + (base + uint32_t(x1) + 1));
    }
  }
  if constexpr (kConcat)
    stream_concat<W>(lo, hi, pts, visit);
  else
    stream_ranges<W>(lo, hi, pts, visit);
}

// scan_block27 restricted to the cells that can hold a point within sqrt(thr)
// of p, for a query inside the grid (hx,hy,hz = its cell): a block row whose
// (y,z) gap exceeds thr is skipped, and each kept row is cut to the cells whose
// x gap still fits (the query's own column always stays). Gaps use the filing
// margin, so no point within sqrt(thr) is ever cut.
template <int W = 4, bool kConcat = false, typename VisitPoint>
__device__ __forceinline__ void scan_block27_within(const GridDesc& g, const uint32_t* __restrict__ cs,
                                                    const float4* __restrict__ pts, float px,
                                                    float py, float pz, int hx, int hy, int hz,
                                                    float thr, VisitPoint&& visit) {
  const float gxl = axis_gap(px, g.ox, g.h, g.margin, hx - 1);
  const float gxr = axis_gap(px, g.ox, g.h, g.margin, hx + 1);
  float gy2[3], gz2[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float gy = axis_gap(py, g.oy, g.h, g.margin, hy - 1 + k);
    const float gz = axis_gap(pz, g.oz, g.h, g.margin, hz - 1 + k);
    gy2[k] = gy * gy;
    gz2[k] = gz * gz;
  }
  uint32_t lo[9], hi[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int cz = hz - 1 + k / 3, cy = hy - 1 + k % 3;
    const float gyz = gz2[k / 3] + gy2[k % 3];
    lo[k] = hi[k] = 0;
    if (cz >= 0 && cz < g.dz && cy >= 0 && cy < g.dy && gyz <= thr) {
      const int xa = (hx > 0 && gyz + gxl * gxl <= thr) ? hx - 1 : hx;
      const int xb = (hx + 1 < g.dx && gyz + gxr * gxr <= thr) ? hx + 1 : hx;
      const uint32_t base = (uint32_t(cz) * uint32_t(g.dy) + uint32_t(cy)) * uint32_t(g.dx);
      lo[k] = ldg(cs + (base + uint32_t(xa)));
      hi[k] = ldg(cs + (base + uint32_t(xb) + 1));
    }
  }
  if constexpr (kConcat)
    stream_concat<W>(lo, hi, pts, visit);
  else
    stream_ranges<W>(lo, hi, pts, visit);
}

// Kahan-compensated f32 accumulator.
struct Kahan {
  float s, c;
  __device__ __forceinline__ void init() { s = 0.0f; c = 0.0f; }
  __device__ __forceinline__ void add(float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

// Error-free transformation: a + b = s + e exactly.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

}  // namespace sr
