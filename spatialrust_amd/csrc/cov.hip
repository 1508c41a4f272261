// Exact k-NN and per-point plane covariances on a device uniform grid.
//
// k-NN path: the GPU replacement of `covariances` + `regularize`
// (crates/spatialrust-registration/src/gicp.rs:259-291): for every point, its
// k nearest neighbours including itself (KdTree::nearest_k semantics,
// crates/spatialrust-search/src/kdtree.rs:306-346), the covariance of that set,
// and the eigenvector n of its smallest eigenvalue. The regularized covariance
// eps*v0v0^T + v1v1^T + v2v2^T equals I - (1-eps) n n^T because the eigenvectors
// are orthonormal, so only n is stored (float4(n, 1)); the iteration kernel
// rebuilds the 3x3 from it.
//
// Search: rings of cells around the query's cell (Chebyshev distance r = 0, 1,
// ...), skipping cells whose box lies farther than the current k-th distance,
// and stopping once every point outside the scanned block is provably farther
// than the k-th neighbour. Neighbours are kept sorted by (d2, index) in
// registers, so ties resolve to the lower index as in brute_force_knn
// (crates/spatialrust-search/src/brute.rs:84-110).
//
// Radius path: COV_GRID_WGSL semantics (crates/spatialrust-gpu/src/kernels/
// covariances_grid.rs:81-206) — cell = radius, 27-cell gather of d2 <= r2,
// f32 two-pass moments, 16 cyclic Jacobi sweeps, eigenvalues (eps, 1, 1),
// count < 3 gives eps*I — evaluated with the same operation order so the
// result matches a CPU evaluation of that kernel bit for bit.

#include <algorithm>

#include "internal.hpp"

namespace sr {

template <int K>
struct KList {
  float d[K];
  uint32_t id[K];   // original index (tie-break key)
  uint32_t pos[K];  // grid position (for re-reading the point)
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      d[j] = INFINITY;
      id[j] = kNone;
      pos[j] = kNone;
    }
  }
  // Bubble the candidate into its sorted place; the last entry falls off.
  __device__ __forceinline__ void insert(float nd, uint32_t ni, uint32_t np) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      bool less = (nd < d[j]) || (nd == d[j] && ni < id[j]);
      float td = less ? d[j] : nd;
      uint32_t ti = less ? id[j] : ni;
      uint32_t tp = less ? pos[j] : np;
      d[j] = less ? nd : d[j];
      id[j] = less ? ni : id[j];
      pos[j] = less ? np : pos[j];
      nd = td;
      ni = ti;
      np = tp;
    }
  }
};

// Exact k nearest neighbours (k <= K) of (px,py,pz) among the grid's points.
// Returns the number found (min(k, npoints)); list entries beyond k are unused.
template <int K>
__device__ int knn_search(const GridDesc& g, const uint32_t* __restrict__ cs,
                          const float4* __restrict__ gp, float px, float py, float pz, int k,
                          KList<K>& L, SearchStats* stats = nullptr) {
  L.init();
  uint32_t n_rings = 0, n_cells = 0, n_pts = 0;
  const int hx = cell_coord_free(px, g.ox, g.inv_h), hy = cell_coord_free(py, g.oy, g.inv_h),
            hz = cell_coord_free(pz, g.oz, g.inv_h);
  const int rmax = max(max(max(hx, g.dx - 1 - hx), max(hy, g.dy - 1 - hy)),
                       max(hz, g.dz - 1 - hz));
  // Rings closer than r0 hold no grid cell (query outside the grid box).
  const int r0 = max(max(max(-hx, hx - (g.dx - 1)), max(-hy, hy - (g.dy - 1))),
                     max(max(-hz, hz - (g.dz - 1)), 0));
  float worst_d = INFINITY;
  uint32_t worst_i = kNone;
  // admission test on the packed (d2, index) key: one u64 compare (the list
  // itself keeps d and id separately, which needs fewer registers)
  uint64_t worst_key = ~0ull;
  auto consider = [&](uint32_t s, const float4& q) {
    const float d2 = dist2(q.x, q.y, q.z, px, py, pz);
    const uint32_t qi = __float_as_uint(q.w);
    const uint64_t key = (uint64_t(__float_as_uint(d2)) << 32) | qi;
    if (key < worst_key && d2 < INFINITY) {  // non-finite distances never enter
      L.insert(d2, qi, s);
#pragma unroll
      for (int j = 0; j < K; ++j) {
        if (j == k - 1) {
          worst_d = L.d[j];
          worst_i = L.id[j];
        }
      }
      worst_key = worst_i == kNone ? ~0ull : (uint64_t(__float_as_uint(worst_d)) << 32) | worst_i;
    }
  };
  int rfirst = r0;
  if (r0 == 0) {
    // Rings 0 and 1 as nine contiguous point ranges, then rings >= 2 only if
    // the k-th distance is not yet certified by the block's clearance.
    ++n_rings;
    scan_block27(g, cs, gp, hx, hy, hz, [&](uint32_t s, const float4& q) {
      ++n_pts;
      consider(s, q);
    });
    if (worst_i != kNone) {
      const float cl = ring_clearance(g, px, py, pz, hx, hy, hz, 1);
      if (cl > 0.0f && worst_d < cl * cl) {
        stats_add(stats, n_rings, n_cells, n_pts, kBlockDone);
        return min(k, int(g.npoints));
      }
    }
    rfirst = 2;
  }
  for (int r = rfirst; r <= rmax; ++r) {
    ++n_rings;
    scan_ring<1>(g, cs, gp, px, py, pz, hx, hy, hz, r, worst_d, n_cells, [&](uint32_t s, const float4& q) {
      ++n_pts;
      consider(s, q);
    });
    if (worst_i != kNone) {
      const float cl = ring_clearance(g, px, py, pz, hx, hy, hz, r);
      if (cl > 0.0f && worst_d < cl * cl) break;
    }
  }
  stats_add(stats, n_rings, n_cells, n_pts, kRingFound);
  return min(k, int(g.npoints));
}

// Smallest-eigenvalue eigenvector of a symmetric 3x3 (cyclic Jacobi, f32).
__device__ __forceinline__ void smallest_eigvec(float a00, float a11, float a22, float a01,
                                                float a02, float a12, float n[3]) {
  float a[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
  float v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const float scale = fabsf(a00) + fabsf(a11) + fabsf(a22) + 1e-30f;
  for (int sweep = 0; sweep < 10; ++sweep) {
    float off = fabsf(a[0][1]) + fabsf(a[0][2]) + fabsf(a[1][2]);
    if (off <= 1e-9f * scale) break;
#pragma unroll
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      float apq = a[p][q];
      if (fabsf(apq) < 1e-30f) continue;
      float phi = 0.5f * (a[q][q] - a[p][p]) / apq;
      float t = (phi >= 0.0f ? 1.0f : -1.0f) / (fabsf(phi) + sqrtf(1.0f + phi * phi));
      float c = 1.0f / sqrtf(1.0f + t * t), s = t * c;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        float arp = a[r][p], arq = a[r][q];
        a[r][p] = c * arp - s * arq;
        a[r][q] = s * arp + c * arq;
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        float apr = a[p][r], aqr = a[q][r];
        a[p][r] = c * apr - s * aqr;
        a[q][r] = s * apr + c * aqr;
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        float vrp = v[r][p], vrq = v[r][q];
        v[r][p] = c * vrp - s * vrq;
        v[r][q] = s * vrp + c * vrq;
      }
    }
  }
  int mi = 0;
  if (a[1][1] < a[mi][mi]) mi = 1;
  if (a[2][2] < a[mi][mi]) mi = 2;
  float x = v[0][mi], y = v[1][mi], z = v[2][mi];
  float len = sqrtf(x * x + y * y + z * z);
  float inv = len > 0.0f ? 1.0f / len : 0.0f;
  n[0] = x * inv;
  n[1] = y * inv;
  n[2] = z * inv;
}

// Per-lane k-th distance by a sorting network: d[0..K] holds the K+1
// smallest squared distances seen so far, ascending; a candidate is merged by
// min/max exchanges (two VALU per slot, no branch), so a wave pays the same
// for every candidate whichever lanes would have inserted it.
template <int K>
struct KDist {
  float d[K + 1];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int j = 0; j <= K; ++j) d[j] = INFINITY;
  }
  // Inserting x into the sorted d: slot j becomes min(d[j], max(d[j-1], x)),
  // the median of (d[j-1], d[j], x) since d[j-1] <= d[j] (slot 0: the median of
  // (d[0], x, -inf), their minimum). One v_med3_f32 per slot, every slot
  // independent of the others (no chain through x), written from the top so
  // each reads the old d[j-1]. The distances are >= +0 or +inf (NaN is mapped
  // to +inf), so the median is the ordered one.
  __device__ __forceinline__ void add(float x) {
    x = x < INFINITY ? x : INFINITY;  // NaN sorts as +inf
#pragma unroll
    for (int j = K; j >= 1; --j) d[j] = __builtin_amdgcn_fmed3f(d[j - 1], d[j], x);
    d[0] = __builtin_amdgcn_fmed3f(d[0], x, -INFINITY);
  }
  __device__ __forceinline__ float at(int j) const {  // d[j] for a runtime j <= K
    float v = d[0];
#pragma unroll
    for (int m = 1; m <= K; ++m)
      if (m == j) v = d[m];
    return v;
  }
};

// The same network over u32 indices (the tie-break among equal distances).
template <int K>
struct KIds {
  uint32_t d[K + 1];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int j = 0; j <= K; ++j) d[j] = 0xFFFFFFFFu;
  }
  __device__ __forceinline__ void add(uint32_t x) {
#pragma unroll
    for (int j = 0; j <= K; ++j) {
      const uint32_t lo = min(d[j], x);
      x = max(d[j], x);
      d[j] = lo;
    }
  }
  __device__ __forceinline__ uint32_t at(int j) const {
    uint32_t v = d[0];
#pragma unroll
    for (int m = 1; m <= K; ++m)
      if (m == j) v = d[m];
    return v;
  }
};

// Slot j of a network: a fixed register when k equals the network size K (j is
// then the compile-time jfix), else a runtime pick. A runtime index into the
// register array would otherwise be served through per-lane scratch memory
// (the whole network stored at each read).
template <bool kExact, int K>
__device__ __forceinline__ float kdist_slot(const KDist<K>& D, int j, int jfix) {
  if constexpr (kExact) return D.d[jfix];
  else return D.at(j);
}

// The 3x3x3 block of a k-NN pass is streamed as one sequence of its nine
// ranges, kKnnBlockW loads in flight: one lane's chain of dependent loads is
// then the block's points / kKnnBlockW steps instead of a few steps per range.
#ifndef SR_KNN_BLOCK_W
#define SR_KNN_BLOCK_W 8
#endif
constexpr int kKnnBlockW = SR_KNN_BLOCK_W;
#ifndef SR_KNN_CONCAT
#define SR_KNN_CONCAT 1
#endif
constexpr bool kKnnConcat = SR_KNN_CONCAT != 0;  // one stream over the nine ranges


// The candidates of an inside query: the 3x3x3 block, then rings 2..rend cut
// to radius^2 thr (a superset of the points within sqrt(thr)).
template <typename Visit>
__device__ __forceinline__ void knn_candidates(const GridDesc& g, const uint32_t* __restrict__ cs,
                                               const float4* __restrict__ gp, float px, float py,
                                               float pz, int hx, int hy, int hz, int rend,
                                               const float& thr, Visit&& visit) {
  scan_block27<kKnnBlockW, kKnnConcat>(g, cs, gp, hx, hy, hz, visit);
  uint32_t rows = 0;
  for (int r = 2; r <= rend; ++r) scan_ring<1>(g, cs, gp, px, py, pz, hx, hy, hz, r, thr, rows, visit);
}

// k-NN normals. The k nearest (ties to the lower index) are found in two
// passes over the same candidates instead of keeping a sorted neighbour list:
// (1) the k-th and (k+1)-th smallest distances by the sorting network, the
// rings extended until the k-th (tau) is certified; (2) the first and second
// moments, relative to the query, of the points closer than tau plus those at
// tau with the lowest indices (exactly k). Only when the k-th and (k+1)-th
// distances are equal does a pass over the points at distance tau pick the
// index cut.
//
// A point still uncertified after kKnnLaneRings rings (an isolated point of a
// sparse cloud, whose neighbours may lie dozens of rings away) is queued for
// k_knn_normals_far instead, so one such point does not hold its whole wave.
constexpr int kKnnLaneRings = 4;

// Output of one point: its grid-order record (point and normal); optionally
// the regularized covariance and the normal in original order.
__device__ __forceinline__ void write_knn_normal(uint32_t i, const float4& p, const float n[3],
                                                 float b, float4* __restrict__ rec,
                                                 float* __restrict__ cov6,
                                                 float* __restrict__ normal3,
                                                 float4* __restrict__ nrm_orig = nullptr,
                                                 float* __restrict__ tau_orig = nullptr,
                                                 float tau = 0.0f) {
  const uint32_t orig = __float_as_uint(p.w);
  rec[kRec * size_t(i)] = p;
  rec[kRec * size_t(i) + 1] = make_float4(n[0], n[1], n[2], 1.0f);
  if (nrm_orig) nrm_orig[orig] = make_float4(n[0], n[1], n[2], 1.0f);
  if (tau_orig) tau_orig[orig] = tau;
  if (cov6) {
    float* o = cov6 + 6 * size_t(orig);
    o[0] = 1.0f - b * n[0] * n[0];
    o[1] = 1.0f - b * n[1] * n[1];
    o[2] = 1.0f - b * n[2] * n[2];
    o[3] = -b * n[0] * n[1];
    o[4] = -b * n[0] * n[2];
    o[5] = -b * n[1] * n[2];
  }
  if (normal3) {
    normal3[3 * size_t(orig) + 0] = n[0];
    normal3[3 * size_t(orig) + 1] = n[1];
    normal3[3 * size_t(orig) + 2] = n[2];
  }
}

// The cloud of launch block b: the last cloud whose first block is <= b.
__device__ __forceinline__ uint32_t knn_cloud_of_block(const KnnCloud* __restrict__ clouds,
                                                       uint32_t nclouds, uint32_t b) {
  uint32_t lo = 0, hi = nclouds;  // clouds[lo].block_begin <= b < clouds[hi].block_begin
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (clouds[mid].block_begin <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

#ifndef SR_KNN_WAVES
#define SR_KNN_WAVES 5
#endif
// Networks above 16 entries hold too many registers for 5 waves per SIMD (41
// VGPRs spilled at K = 20): they are compiled for 4 (256 pairs at k = 20:
// +1.5% batch throughput; k = 15 stays 7% faster at 5).
template <int K>
constexpr int knn_waves() { return K > 16 ? (SR_KNN_WAVES < 4 ? SR_KNN_WAVES : 4) : SR_KNN_WAVES; }

template <int K, bool kExact>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(knn_waves<K>()))) k_knn_normals(const KnnCloud* __restrict__ clouds,
                                                        uint32_t nclouds, int k,
                                                        float one_minus_eps,
                                                        unsigned long long* __restrict__ far_queue,
                                                        uint32_t* __restrict__ far_count,
                                                        SearchStats* __restrict__ stats) {
  const uint32_t c = knn_cloud_of_block(clouds, nclouds, blockIdx.x);
  const KnnCloud& C = clouds[c];
  const GridDesc g = C.g;
  const uint32_t* __restrict__ cs = C.cs;
  const float4* __restrict__ gp = C.gp;
  float4* __restrict__ rec = C.rec;
  float* __restrict__ cov6 = C.cov6;
  float* __restrict__ normal3 = C.normal3;
  uint32_t i = (blockIdx.x - C.block_begin) * blockDim.x + threadIdx.x;
  if (C.subset) {
    if (i >= *C.subset_n) return;
    i = C.subset[i];
  }
  if (i >= g.npoints) return;
  const float4 p = gp[i];
  const int hx = cell_coord_free(p.x, g.ox, g.inv_h), hy = cell_coord_free(p.y, g.oy, g.inv_h),
            hz = cell_coord_free(p.z, g.oz, g.inv_h);
  const int rmax = max(max(max(hx, g.dx - 1 - hx), max(hy, g.dy - 1 - hy)),
                       max(hz, g.dz - 1 - hz));
  // pass 1
  float tau, next;
  int c_lt = 0, rend = 1;
  {
    KDist<K> D;
    D.init();
    auto net = [&](uint32_t, const float4& q) { D.add(dist2(q.x, q.y, q.z, p.x, p.y, p.z)); };
    scan_block27<kKnnBlockW, kKnnConcat>(g, cs, gp, hx, hy, hz, net);
    float worst = kdist_slot<kExact>(D, k - 1, K - 1);
    const float cl = ring_clearance(g, p.x, p.y, p.z, hx, hy, hz, 1);
    if (!(worst < INFINITY && cl > 0.0f && worst < cl * cl)) {
      uint32_t rows = 0;
      bool certified = false;
      for (int r = 2; r <= min(rmax, kKnnLaneRings); ++r) {
        scan_ring<1>(g, cs, gp, p.x, p.y, p.z, hx, hy, hz, r, worst, rows, net);
        rend = r;
        worst = kdist_slot<kExact>(D, k - 1, K - 1);
        const float c = ring_clearance(g, p.x, p.y, p.z, hx, hy, hz, r);
        if (worst < INFINITY && c > 0.0f && worst < c * c) {
          certified = true;
          break;
        }
      }
      if (!certified && rmax > kKnnLaneRings) {
        far_queue[atomicAdd(far_count, 1u)] = (uint64_t(c) << 32) | i;
        return;
      }
    }
    tau = worst;
    next = kdist_slot<kExact>(D, k, K);
#pragma unroll
    for (int j = 0; j < K; ++j) c_lt += (j < k && D.d[j] < tau) ? 1 : 0;
  }
  // search statistics: rings per query (also as "cells", so the maximum is kept)
  if (stats) stats_add(stats, uint32_t(rend), uint32_t(rend), 0, rend == 1 ? kBlockDone : kRingFound);
  uint32_t id_cut = 0xFFFFFFFFu;
  if (tau < INFINITY && next == tau) {
    // k-th and (k+1)-th tie: of the points at distance tau, take the
    // k - c_lt lowest indices
    KIds<K> T;
    T.init();
    knn_candidates(g, cs, gp, p.x, p.y, p.z, hx, hy, hz, rend, tau, [&](uint32_t, const float4& q) {
      if (dist2(q.x, q.y, q.z, p.x, p.y, p.z) == tau) T.add(__float_as_uint(q.w));
    });
    id_cut = T.at(k - c_lt - 1);
  }
  auto selected = [&](const float4& q) {
    const float d2 = dist2(q.x, q.y, q.z, p.x, p.y, p.z);
    return d2 < tau || (d2 == tau && __float_as_uint(q.w) <= id_cut);
  };
  float n[3] = {1.0f, 0.0f, 0.0f};  // count < 2: identity -> eigenvector e_x (gicp.rs:268-270)
  float mx = 0.0f, my = 0.0f, mz = 0.0f, c00 = 0, c11 = 0, c22 = 0, c01 = 0, c02 = 0, c12 = 0;
  int found = 0;
  // one pass: first and second moments relative to the query, then
  // C = S2 - S1 S1^T / n (the offsets are within the k-NN radius, so the
  // subtraction loses little)
  knn_candidates(g, cs, gp, p.x, p.y, p.z, hx, hy, hz, rend, tau, [&](uint32_t, const float4& q) {
    if (selected(q)) {
      const float ex = q.x - p.x, ey = q.y - p.y, ez = q.z - p.z;
      mx += ex;
      my += ey;
      mz += ez;
      c00 += ex * ex;
      c11 += ey * ey;
      c22 += ez * ez;
      c01 += ex * ey;
      c02 += ex * ez;
      c12 += ey * ez;
      ++found;
    }
  });
  if (found >= 2) {
    const float invn = 1.0f / float(found);
    c00 -= mx * mx * invn;
    c11 -= my * my * invn;
    c22 -= mz * mz * invn;
    c01 -= mx * my * invn;
    c02 -= mx * mz * invn;
    c12 -= my * mz * invn;
  }

  if (found >= 2) smallest_eigvec(c00, c11, c22, c01, c02, c12, n);
  write_knn_normal(i, p, n, one_minus_eps, rec, cov6, normal3, C.nrm_orig,
                   C.tau_orig, tau);
}

// ---- far points: one wave per queued point --------------------------------
//
// The same three passes with the point's rings shared by the 64 lanes: ring
// r's (y,z) rows (face rows whole, interior rows their two end cells) are dealt
// round robin, each lane keeps its own sorting network, and the wave merges
// them only where needed. After each ring the wave counts its candidates
// closer than the ring's clearance (each lane's count is exact up to K+1): k of
// them certify the k-th distance. The rings are cut to the smallest of the
// lanes' k-th distances, which bounds the wave's k-th from above.

// Visits this lane's share of the points filed in Chebyshev ring r (r >= 0)
// whose cells can hold a point within sqrt(thr) of p.
template <typename Visit>
__device__ __forceinline__ void coop_scan_ring(const GridDesc& g, const uint32_t* __restrict__ cs,
                                               const float4* __restrict__ gp, float px, float py,
                                               float pz, int hx, int hy, int hz, int r, float thr,
                                               Visit&& visit) {
  const int lane = int(threadIdx.x & (kWave - 1));
  const int side = 2 * r + 1;
  for (int t = lane; t < side * side; t += kWave) {
    const int zi = t / side, yi = t - zi * side;
    const int cz = hz - r + zi, cy = hy - r + yi;
    if (cz < 0 || cz >= g.dz || cy < 0 || cy >= g.dy) continue;
    if (!((g.rows[size_t(cz) * g.wps + uint32_t(cy >> 5)] >> (cy & 31)) & 1u)) continue;
    const float gz = axis_gap(pz, g.oz, g.h, g.margin, cz);
    const float gy = axis_gap(py, g.oy, g.h, g.margin, cy);
    const float gyz = gz * gz + gy * gy;
    if (gyz > thr) continue;
    const uint32_t cbase = (uint32_t(cz) * uint32_t(g.dy) + uint32_t(cy)) * uint32_t(g.dx);
    uint32_t lo[2] = {0, 0}, hi[2] = {0, 0};
    if (zi == 0 || zi == side - 1 || yi == 0 || yi == side - 1) {
      int x0 = max(hx - r, 0), x1 = min(hx + r, g.dx - 1);
      if (thr < INFINITY) {
        const float rx = __builtin_amdgcn_sqrtf(thr - gyz) + g.margin;
        x0 = max(x0, cell_coord_free(px - rx, g.ox, g.inv_h) - 1);
        x1 = min(x1, cell_coord_free(px + rx, g.ox, g.inv_h) + 1);
      }
      if (x0 <= x1) {
        lo[0] = cs[cbase + uint32_t(x0)];
        hi[0] = cs[cbase + uint32_t(x1) + 1];
      }
    } else {
      // interior row of the ring: its two end cells
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int cx = e == 0 ? hx - r : hx + r;
        if (cx < 0 || cx >= g.dx) continue;
        const float gx = axis_gap(px, g.ox, g.h, g.margin, cx);
        if (gyz + gx * gx > thr) continue;
        lo[e] = cs[cbase + uint32_t(cx)];
        hi[e] = cs[cbase + uint32_t(cx) + 1];
      }
    }
    stream_ranges<1>(lo, hi, gp, visit);
  }
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

template <int K, bool kExact>
__global__ void __launch_bounds__(kBlock) k_knn_normals_far(const KnnCloud* __restrict__ clouds,
                                                            int k, float one_minus_eps,
                                                            const unsigned long long* __restrict__ far_queue,
                                                            const uint32_t* __restrict__ far_count) {
  const uint32_t n_far = *far_count;
  const uint32_t waves = gridDim.x * (kBlock / kWave);
  const int lane = int(threadIdx.x & (kWave - 1));
  for (uint32_t w = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; w < n_far; w += waves) {
    const unsigned long long e = far_queue[w];
    const KnnCloud& C = clouds[uint32_t(e >> 32)];
    const GridDesc g = C.g;
    const uint32_t* __restrict__ cs = C.cs;
    const float4* __restrict__ gp = C.gp;
    const uint32_t i = uint32_t(e);
    const float4 p = gp[i];
    const int hx = cell_coord_free(p.x, g.ox, g.inv_h), hy = cell_coord_free(p.y, g.oy, g.inv_h),
              hz = cell_coord_free(p.z, g.oz, g.inv_h);
    const int rmax = max(max(max(hx, g.dx - 1 - hx), max(hy, g.dy - 1 - hy)),
                         max(hz, g.dz - 1 - hz));
    // pass 1: lane networks, rings until k candidates lie inside the clearance
    KDist<K> D;
    D.init();
    auto net = [&](uint32_t, const float4& q) { D.add(dist2(q.x, q.y, q.z, p.x, p.y, p.z)); };
    float thr = INFINITY;
    int rend = 0;
    for (int r = 0; r <= rmax; ++r) {
      coop_scan_ring(g, cs, gp, p.x, p.y, p.z, hx, hy, hz, r, thr, net);
      rend = r;
      const float c = ring_clearance(g, p.x, p.y, p.z, hx, hy, hz, r);
      float inside = 0.0f;
#pragma unroll
      for (int j = 0; j <= K; ++j) inside += (c > 0.0f && D.d[j] < c * c) ? 1.0f : 0.0f;
      if (wave_sum_f(inside) >= float(k)) break;
      float lane_kth = kdist_slot<kExact>(D, k - 1, K - 1);
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) lane_kth = fminf(lane_kth, __shfl_xor(lane_kth, off));
      thr = lane_kth;
    }
    // merge the lane networks: every lane ends with the wave's K+1 smallest
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
      float other[K + 1];
#pragma unroll
      for (int j = 0; j <= K; ++j) other[j] = __shfl_xor(D.d[j], off);
#pragma unroll
      for (int j = 0; j <= K; ++j) D.add(other[j]);
    }
    const float tau = kdist_slot<kExact>(D, k - 1, K - 1), next = kdist_slot<kExact>(D, k, K);
    int c_lt = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) c_lt += (j < k && D.d[j] < tau) ? 1 : 0;
    uint32_t id_cut = 0xFFFFFFFFu;
    if (tau < INFINITY && next == tau) {
      KIds<K> T;
      T.init();
      for (int r = 0; r <= rend; ++r)
        coop_scan_ring(g, cs, gp, p.x, p.y, p.z, hx, hy, hz, r, tau, [&](uint32_t, const float4& q) {
          if (dist2(q.x, q.y, q.z, p.x, p.y, p.z) == tau) T.add(__float_as_uint(q.w));
        });
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) {
        uint32_t other[K + 1];
#pragma unroll
        for (int j = 0; j <= K; ++j) other[j] = __shfl_xor(T.d[j], off);
#pragma unroll
        for (int j = 0; j <= K; ++j) T.add(other[j]);
      }
      id_cut = T.at(k - c_lt - 1);
    }
    auto selected = [&](const float4& q) {
      const float d2 = dist2(q.x, q.y, q.z, p.x, p.y, p.z);
      return d2 < tau || (d2 == tau && __float_as_uint(q.w) <= id_cut);
    };
    // pass 2: mean of the selected points, relative to the query
    float mx = 0.0f, my = 0.0f, mz = 0.0f, cnt = 0.0f;
    for (int r = 0; r <= rend; ++r)
      coop_scan_ring(g, cs, gp, p.x, p.y, p.z, hx, hy, hz, r, tau, [&](uint32_t, const float4& q) {
        if (selected(q)) {
          mx += q.x - p.x;
          my += q.y - p.y;
          mz += q.z - p.z;
          cnt += 1.0f;
        }
      });
    mx = wave_sum_f(mx);
    my = wave_sum_f(my);
    mz = wave_sum_f(mz);
    const int found = int(wave_sum_f(cnt));
    float nv[3] = {1.0f, 0.0f, 0.0f};  // count < 2: identity -> eigenvector e_x (gicp.rs:268-270)
    if (found >= 2) {
      const float invn = 1.0f / float(found);
      mx *= invn;
      my *= invn;
      mz *= invn;
      // pass 3: centred sums
      float c00 = 0, c11 = 0, c22 = 0, c01 = 0, c02 = 0, c12 = 0;
      for (int r = 0; r <= rend; ++r)
        coop_scan_ring(g, cs, gp, p.x, p.y, p.z, hx, hy, hz, r, tau, [&](uint32_t, const float4& q) {
          if (selected(q)) {
            const float ex = (q.x - p.x) - mx, ey = (q.y - p.y) - my, ez = (q.z - p.z) - mz;
            c00 += ex * ex;
            c11 += ey * ey;
            c22 += ez * ez;
            c01 += ex * ey;
            c02 += ex * ez;
            c12 += ey * ez;
          }
        });
      c00 = wave_sum_f(c00);
      c11 = wave_sum_f(c11);
      c22 = wave_sum_f(c22);
      c01 = wave_sum_f(c01);
      c02 = wave_sum_f(c02);
      c12 = wave_sum_f(c12);
      smallest_eigvec(c00, c11, c22, c01, c02, c12, nv);
    }
    if (lane == 0)
      write_knn_normal(i, p, nv, one_minus_eps, C.rec, C.cov6, C.normal3,
                       C.nrm_orig, C.tau_orig, tau);
  }
}

template <int K, bool kExact>
static void launch_knn_normals(const std::vector<KnnCloud>& host, int k, float b, hipStream_t s) {
  uint64_t n = 0;
  for (const KnnCloud& c : host) n += c.g.npoints;
  const uint32_t blocks = host.back().block_begin +
                          (host.back().g.npoints + kBlock - 1) / kBlock;
  DevBuf<KnnCloud> clouds(host.size());
  upload(clouds.p, host.data(), sizeof(KnnCloud) * host.size(), s);
  DevBuf<unsigned long long> far_queue(std::max<uint64_t>(n, 1));
  DevBuf<uint32_t> far_count(1);
  fill_u32(far_count.p, 0u, 1, s);
  ProfScope prof("knn_normals", s);
  hipLaunchKernelGGL((k_knn_normals<K, kExact>), dim3(blocks), dim3(kBlock), 0, s, clouds.p,
                     uint32_t(host.size()), k, b, far_queue.p, far_count.p,
                     prof_stats("knn_normals"));
  // one wave per far point, a grid the chip holds at once (few points qualify)
  const uint32_t far_blocks =
      uint32_t(std::min<uint64_t>((n + kBlock / kWave - 1) / (kBlock / kWave), 1024));
  hipLaunchKernelGGL((k_knn_normals_far<K, kExact>), dim3(far_blocks), dim3(kBlock), 0, s, clouds.p, k, b,
                     far_queue.p, far_count.p);
  // the descriptor table and the queue are stream-ordered scratch: the
  // allocator hands them out again only to later work on this stream
}

void knn_normals_many(const std::vector<KnnCloud>& clouds_in, uint32_t k, double eps,
                      hipStream_t s) {
  std::vector<KnnCloud> clouds;
  uint32_t blocks = 0;
  for (KnnCloud c : clouds_in) {
    if (c.g.npoints == 0) continue;
    c.block_begin = blocks;
    blocks += (c.g.npoints + kBlock - 1) / kBlock;
    clouds.push_back(c);
  }
  if (clouds.empty()) return;
  const float b = float(1.0 - eps);
  if (k == 12) launch_knn_normals<12, true>(clouds, int(k), b, s);
  else if (k == 15) launch_knn_normals<15, true>(clouds, int(k), b, s);
  else if (k == 20) launch_knn_normals<20, true>(clouds, int(k), b, s);
  else if (k <= 8) launch_knn_normals<8, false>(clouds, int(k), b, s);
  else if (k <= 32) launch_knn_normals<32, false>(clouds, int(k), b, s);
  else if (k <= 64) launch_knn_normals<64, false>(clouds, int(k), b, s);
  else fail(3, "k_neighbors above 64 is not supported by the HIP covariance kernel");
}

KnnCloud knn_cloud(const Grid& g, float4* rec, float* cov6, float* normal3) {
  KnnCloud c{};
  c.g = g.desc;
  c.cs = g.cell_start.p;
  c.gp = g.pts.p;
  c.rec = rec;
  c.cov6 = cov6;
  c.normal3 = normal3;
  return c;
}

void knn_normals(const Grid& g, uint32_t k, double eps, float4* rec, float* cov6, float* normal3,
                 hipStream_t s) {
  if (g.desc.npoints == 0) return;
  knn_normals_many({knn_cloud(g, rec, cov6, normal3)}, k, eps, s);
}

// ------------------------------------------------------------ k-NN query ----

template <int K>
__global__ void __launch_bounds__(kBlock) k_knn_query(GridDesc g, const uint32_t* __restrict__ cs,
                                                      const float4* __restrict__ gp,
                                                      const float4* __restrict__ qs, uint32_t nq,
                                                      int k, int32_t* __restrict__ out_idx,
                                                      float* __restrict__ out_d2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const float4 q = qs[i];
  KList<K> L;
  const int found = knn_search<K>(g, cs, gp, q.x, q.y, q.z, k, L);
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j < k) {
      out_idx[size_t(i) * k + j] = j < found ? int32_t(L.id[j]) : -1;
      out_d2[size_t(i) * k + j] = j < found ? L.d[j] : -1.0f;
    }
  }
}

void knn_query(const Grid& g, const float4* queries, uint32_t nq, uint32_t k, int32_t* out_idx,
               float* out_d2, hipStream_t s) {
  if (nq == 0 || k == 0) return;
  const dim3 grid((nq + kBlock - 1) / kBlock);
#define SR_KNNQ(KK)                                                                         \
  hipLaunchKernelGGL(k_knn_query<KK>, grid, dim3(kBlock), 0, s, g.desc, g.cell_start.p, g.pts.p, \
                     queries, nq, int(k), out_idx, out_d2)
  if (k == 1) SR_KNNQ(1);
  else if (k <= 8) SR_KNNQ(8);
  else if (k <= 16) SR_KNNQ(16);
  else if (k <= 32) SR_KNNQ(32);
  else if (k <= 64) SR_KNNQ(64);
  else fail(3, "k above 64 is not supported by the HIP k-NN kernel");
#undef SR_KNNQ
}

// ------------------------------------------------------- radius (WGSL) ----

__device__ __forceinline__ void wgsl_rotate(float a[3][3], float v[3][3], int p, int q) {
  float apq = a[p][q];
  if (fabsf(apq) < 1e-20f) return;
  float app = a[p][p], aqq = a[q][q];
  float phi = __fdiv_rn(__fmul_rn(0.5f, __fsub_rn(aqq, app)), apq);
  // sqrtf, not __fsqrt_rn: on gfx950 the latter is not correctly rounded, while
  // sqrtf is under the default correctly-rounded divide/sqrt mode.
  float root = sqrtf(__fadd_rn(1.0f, __fmul_rn(phi, phi)));
  float t = phi >= 0.0f ? __fdiv_rn(1.0f, __fadd_rn(phi, root))
                        : __fdiv_rn(-1.0f, __fadd_rn(-phi, root));
  float c = __fdiv_rn(1.0f, sqrtf(__fadd_rn(1.0f, __fmul_rn(t, t))));
  float s = __fmul_rn(t, c);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float arp = a[r][p], arq = a[r][q];
    a[r][p] = __fsub_rn(__fmul_rn(c, arp), __fmul_rn(s, arq));
    a[r][q] = __fadd_rn(__fmul_rn(s, arp), __fmul_rn(c, arq));
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float apr = a[p][r], aqr = a[q][r];
    a[p][r] = __fsub_rn(__fmul_rn(c, apr), __fmul_rn(s, aqr));
    a[q][r] = __fadd_rn(__fmul_rn(s, apr), __fmul_rn(c, aqr));
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float vrp = v[r][p], vrq = v[r][q];
    v[r][p] = __fsub_rn(__fmul_rn(c, vrp), __fmul_rn(s, vrq));
    v[r][q] = __fadd_rn(__fmul_rn(s, vrp), __fmul_rn(c, vrq));
  }
}

// Two-pass f32 covariance of the radius neighbourhood of p over a grid whose
// edge is the radius, then 16 Jacobi sweeps, in the WGSL kernels' order
// (covariances_grid.rs:81-206, normals_grid.rs:86-160 are identical up to this
// point). Neighbours are read from the grid-ordered copy (ascending index
// inside each cell, the order build_grid produces), so the f32 sums happen in
// the reference's order. Returns the neighbour count; a and v hold the rotated
// matrix and the eigenvectors (columns) when the count is at least 3.
template <typename Range, typename Point>
__device__ __forceinline__ float radius_eigen_with(const GridDesc& g, const float4& p, float r2,
                                                   Range&& range, Point&& point, float (&a)[3][3],
                                                   float (&v)[3][3]) {
  const int cx = cell_coord(p.x, g.ox, g.inv_h, g.dx), cy = cell_coord(p.y, g.oy, g.inv_h, g.dy),
            cz = cell_coord(p.z, g.oz, g.inv_h, g.dz);
  float mx = 0.0f, my = 0.0f, mz = 0.0f, count = 0.0f;
  for (int dz = -1; dz <= 1; ++dz) {
    int nz = cz + dz;
    if (nz < 0 || nz >= g.dz) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      int ny = cy + dy;
      if (ny < 0 || ny >= g.dy) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        int nx = cx + dx;
        if (nx < 0 || nx >= g.dx) continue;
        uint32_t lo, hi;
        range(nx, ny, nz, lo, hi);
        for (uint32_t s = lo; s < hi; ++s) {
          const float4 q = point(s);
          float ddx = __fsub_rn(q.x, p.x), ddy = __fsub_rn(q.y, p.y), ddz = __fsub_rn(q.z, p.z);
          float d2 = __fadd_rn(__fadd_rn(__fmul_rn(ddx, ddx), __fmul_rn(ddy, ddy)), __fmul_rn(ddz, ddz));
          if (d2 <= r2) {
            mx = __fadd_rn(mx, q.x);
            my = __fadd_rn(my, q.y);
            mz = __fadd_rn(mz, q.z);
            count = __fadd_rn(count, 1.0f);
          }
        }
      }
    }
  }
  if (count < 3.0f) return count;
  mx = __fdiv_rn(mx, count);
  my = __fdiv_rn(my, count);
  mz = __fdiv_rn(mz, count);
  float c00 = 0, c11 = 0, c22 = 0, c01 = 0, c02 = 0, c12 = 0;
  for (int dz = -1; dz <= 1; ++dz) {
    int nz = cz + dz;
    if (nz < 0 || nz >= g.dz) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      int ny = cy + dy;
      if (ny < 0 || ny >= g.dy) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        int nx = cx + dx;
        if (nx < 0 || nx >= g.dx) continue;
        uint32_t lo, hi;
        range(nx, ny, nz, lo, hi);
        for (uint32_t s = lo; s < hi; ++s) {
          const float4 q = point(s);
          float rx = __fsub_rn(q.x, p.x), ry = __fsub_rn(q.y, p.y), rz = __fsub_rn(q.z, p.z);
          float d2 = __fadd_rn(__fadd_rn(__fmul_rn(rx, rx), __fmul_rn(ry, ry)), __fmul_rn(rz, rz));
          if (d2 <= r2) {
            float ex = __fsub_rn(q.x, mx), ey = __fsub_rn(q.y, my), ez = __fsub_rn(q.z, mz);
            c00 = __fadd_rn(c00, __fmul_rn(ex, ex));
            c11 = __fadd_rn(c11, __fmul_rn(ey, ey));
            c22 = __fadd_rn(c22, __fmul_rn(ez, ez));
            c01 = __fadd_rn(c01, __fmul_rn(ex, ey));
            c02 = __fadd_rn(c02, __fmul_rn(ex, ez));
            c12 = __fadd_rn(c12, __fmul_rn(ey, ez));
          }
        }
      }
    }
  }
  a[0][0] = c00; a[0][1] = c01; a[0][2] = c02;
  a[1][0] = c01; a[1][1] = c11; a[1][2] = c12;
  a[2][0] = c02; a[2][1] = c12; a[2][2] = c22;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) v[r][c] = r == c ? 1.0f : 0.0f;
  for (int sweep = 0; sweep < 16; ++sweep) {
    wgsl_rotate(a, v, 0, 1);
    wgsl_rotate(a, v, 0, 2);
    wgsl_rotate(a, v, 1, 2);
  }
  return count;
}

// The same evaluation with the 27-cell neighbourhood read as nine row
// ranges: in grid order the cells x-1 .. x+1 of one (y, z) row are one
// contiguous run of points, visited in the per-cell order (dx ascending, then
// the points ascending), so the f32 sums, and the results, are those of
// radius_eigen_with. The nine ranges' cell starts are loaded together once
// (18 words) and serve both moment passes, and each range is streamed four
// points per round, so a lane waits on a few rounds of loads instead of a
// chain of per-cell and per-point round trips.
constexpr int kRadiusStream = 4;

template <typename Point, typename Pass>
__device__ __forceinline__ void radius_rows_visit(Point&& point, const uint32_t (&lo)[9],
                                                  const uint32_t (&hi)[9], Pass&& visit) {
#pragma unroll 1
  for (int r = 0; r < 9; ++r) {
    const uint32_t e = hi[r];
#pragma unroll 1
    for (uint32_t s = lo[r]; s < e; s += kRadiusStream) {
      float4 q[kRadiusStream];
#pragma unroll
      for (int u = 0; u < kRadiusStream; ++u) q[u] = point(min(s + uint32_t(u), e - 1));
#pragma unroll
      for (int u = 0; u < kRadiusStream; ++u)
        if (s + uint32_t(u) < e) visit(q[u]);
    }
  }
}

// The point's cell and, for each of the nine (y, z) rows around it (r = (dz +
// 1) * 3 + (dy + 1)), its x-cell range [x0, x1); rows outside the grid get
// an empty range (in = false).
__device__ __forceinline__ void radius_rows(const GridDesc& g, const float4& p, int& cy, int& cz,
                                            uint32_t& x0, uint32_t& x1) {
  const int cx = cell_coord(p.x, g.ox, g.inv_h, g.dx);
  cy = cell_coord(p.y, g.oy, g.inv_h, g.dy);
  cz = cell_coord(p.z, g.oz, g.inv_h, g.dz);
  x0 = uint32_t(max(cx - 1, 0));
  x1 = uint32_t(min(cx + 1, g.dx - 1)) + 1u;
}

template <typename Point>
__device__ __forceinline__ float radius_eigen_ranges(const float4& p, float r2,
                                                     const uint32_t (&lo)[9],
                                                     const uint32_t (&hi)[9], Point&& point,
                                                     float (&a)[3][3], float (&v)[3][3]);

__device__ __forceinline__ float radius_eigen_rows(const GridDesc& g,
                                                   const uint32_t* __restrict__ cs,
                                                   const float4* __restrict__ gp, const float4& p,
                                                   float r2, float (&a)[3][3], float (&v)[3][3]) {
  int cy, cz;
  uint32_t x0, x1;
  radius_rows(g, p, cy, cz, x0, x1);
  uint32_t lo[9], hi[9];
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int ny = cy + r % 3 - 1, nz = cz + r / 3 - 1;
    const bool in = ny >= 0 && ny < g.dy && nz >= 0 && nz < g.dz;
    const uint32_t base = in ? (uint32_t(nz) * uint32_t(g.dy) + uint32_t(ny)) * uint32_t(g.dx) : 0u;
    lo[r] = in ? cs[base + x0] : 0u;
    hi[r] = in ? cs[base + x1] : 0u;
  }
  return radius_eigen_ranges(p, r2, lo, hi, [&](uint32_t s) { return gp[s]; }, a, v);
}

// The two moment passes and the Jacobi sweeps over the nine row ranges.
template <typename Point>
__device__ __forceinline__ float radius_eigen_ranges(const float4& p, float r2,
                                                     const uint32_t (&lo)[9],
                                                     const uint32_t (&hi)[9], Point&& point,
                                                     float (&a)[3][3], float (&v)[3][3]) {
  float mx = 0.0f, my = 0.0f, mz = 0.0f, count = 0.0f;
  radius_rows_visit(point, lo, hi, [&](const float4& q) {
    float ddx = __fsub_rn(q.x, p.x), ddy = __fsub_rn(q.y, p.y), ddz = __fsub_rn(q.z, p.z);
    float d2 = __fadd_rn(__fadd_rn(__fmul_rn(ddx, ddx), __fmul_rn(ddy, ddy)), __fmul_rn(ddz, ddz));
    if (d2 <= r2) {
      mx = __fadd_rn(mx, q.x);
      my = __fadd_rn(my, q.y);
      mz = __fadd_rn(mz, q.z);
      count = __fadd_rn(count, 1.0f);
    }
  });
  if (count < 3.0f) return count;
  mx = __fdiv_rn(mx, count);
  my = __fdiv_rn(my, count);
  mz = __fdiv_rn(mz, count);
  float c00 = 0, c11 = 0, c22 = 0, c01 = 0, c02 = 0, c12 = 0;
  radius_rows_visit(point, lo, hi, [&](const float4& q) {
    float rx = __fsub_rn(q.x, p.x), ry = __fsub_rn(q.y, p.y), rz = __fsub_rn(q.z, p.z);
    float d2 = __fadd_rn(__fadd_rn(__fmul_rn(rx, rx), __fmul_rn(ry, ry)), __fmul_rn(rz, rz));
    if (d2 <= r2) {
      float ex = __fsub_rn(q.x, mx), ey = __fsub_rn(q.y, my), ez = __fsub_rn(q.z, mz);
      c00 = __fadd_rn(c00, __fmul_rn(ex, ex));
      c11 = __fadd_rn(c11, __fmul_rn(ey, ey));
      c22 = __fadd_rn(c22, __fmul_rn(ez, ez));
      c01 = __fadd_rn(c01, __fmul_rn(ex, ey));
      c02 = __fadd_rn(c02, __fmul_rn(ex, ez));
      c12 = __fadd_rn(c12, __fmul_rn(ey, ez));
    }
  });
  a[0][0] = c00; a[0][1] = c01; a[0][2] = c02;
  a[1][0] = c01; a[1][1] = c11; a[1][2] = c12;
  a[2][0] = c02; a[2][1] = c12; a[2][2] = c22;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) v[r][c] = r == c ? 1.0f : 0.0f;
  for (int sweep = 0; sweep < 16; ++sweep) {
    wgsl_rotate(a, v, 0, 1);
    wgsl_rotate(a, v, 0, 2);
    wgsl_rotate(a, v, 1, 2);
  }
  return count;
}

// The neighbourhood read straight from the grid in device memory, cell by
// cell (the tiled form's fallback for tiles whose halo exceeds the LDS).
__device__ __forceinline__ float radius_eigen(const GridDesc& g, const uint32_t* __restrict__ cs,
                                              const float4* __restrict__ gp, const float4& p,
                                              float r2, float (&a)[3][3], float (&v)[3][3]) {
  return radius_eigen_with(
      g, p, r2,
      [&](int nx, int ny, int nz, uint32_t& lo, uint32_t& hi) {
        const uint32_t c = (uint32_t(nz) * uint32_t(g.dy) + uint32_t(ny)) * uint32_t(g.dx) + uint32_t(nx);
        lo = cs[c];
        hi = cs[c + 1];
      },
      [&](uint32_t s) { return gp[s]; }, a, v);
}

// Per-point outputs of the radius covariance (COV_GRID_WGSL) path.
__device__ __forceinline__ void radius_cov_out(const float4& p, float count, const float (&a)[3][3],
                                               const float (&v)[3][3], float eps,
                                               float4* __restrict__ nrm_orig,
                                               float* __restrict__ cov6,
                                               float* __restrict__ normal3) {
  const uint32_t orig = __float_as_uint(p.w);
  float reg[6] = {eps, eps, eps, 0.0f, 0.0f, 0.0f};
  float nv[3] = {0.0f, 0.0f, 0.0f};
  float a_iso = eps;  // C = a*I - (1-eps) n n^T; eps*I when count < 3
  if (count >= 3.0f) {
    int mi = 0;
    if (a[1][1] < a[mi][mi]) mi = 1;
    if (a[2][2] < a[mi][mi]) mi = 2;
    for (int k = 0; k < 6; ++k) reg[k] = 0.0f;
    for (int col = 0; col < 3; ++col) {
      float lam = col == mi ? eps : 1.0f;
      float ax = v[0][col], ay = v[1][col], az = v[2][col];
      reg[0] = __fadd_rn(reg[0], __fmul_rn(__fmul_rn(lam, ax), ax));
      reg[1] = __fadd_rn(reg[1], __fmul_rn(__fmul_rn(lam, ay), ay));
      reg[2] = __fadd_rn(reg[2], __fmul_rn(__fmul_rn(lam, az), az));
      reg[3] = __fadd_rn(reg[3], __fmul_rn(__fmul_rn(lam, ax), ay));
      reg[4] = __fadd_rn(reg[4], __fmul_rn(__fmul_rn(lam, ax), az));
      reg[5] = __fadd_rn(reg[5], __fmul_rn(__fmul_rn(lam, ay), az));
    }
    nv[0] = v[0][mi];
    nv[1] = v[1][mi];
    nv[2] = v[2][mi];
    a_iso = 1.0f;
  }
  if (nrm_orig) nrm_orig[orig] = make_float4(nv[0], nv[1], nv[2], a_iso);
  if (cov6) {
#pragma unroll
    for (int k = 0; k < 6; ++k) cov6[6 * size_t(orig) + k] = reg[k];
  }
  if (normal3) {
    normal3[3 * size_t(orig) + 0] = nv[0];
    normal3[3 * size_t(orig) + 1] = nv[1];
    normal3[3 * size_t(orig) + 2] = nv[2];
  }
}

// Per-point outputs of the grid normal (NORMALS_GRID_WGSL) path, oriented to
// the viewpoint when one is given.
__device__ __forceinline__ void grid_normal_out(const float4& p, float count, const float (&a)[3][3],
                                                const float (&v)[3][3], int has_vp, float vx,
                                                float vy, float vz, float4* __restrict__ out4) {
  const uint32_t orig = __float_as_uint(p.w);
  float n0 = 0.0f, n1 = 0.0f, n2 = 1.0f, curv = 0.0f;
  if (count >= 3.0f) {
    const float eig[3] = {a[0][0], a[1][1], a[2][2]};
    int mi = 0;
    if (eig[1] < eig[mi]) mi = 1;
    if (eig[2] < eig[mi]) mi = 2;
    n0 = v[0][mi];
    n1 = v[1][mi];
    n2 = v[2][mi];
    const float len = fmaxf(
        sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(n0, n0), __fmul_rn(n1, n1)), __fmul_rn(n2, n2))), 1e-20f);
    n0 = __fdiv_rn(n0, len);
    n1 = __fdiv_rn(n1, len);
    n2 = __fdiv_rn(n2, len);
    const float tr = __fadd_rn(__fadd_rn(eig[0], eig[1]), eig[2]);
    if (tr > 0.0f) curv = __fdiv_rn(eig[mi], tr);
  }
  if (has_vp) {
    const float wx = __fsub_rn(vx, p.x), wy = __fsub_rn(vy, p.y), wz = __fsub_rn(vz, p.z);
    if (__fadd_rn(__fadd_rn(__fmul_rn(n0, wx), __fmul_rn(n1, wy)), __fmul_rn(n2, wz)) < 0.0f) {
      n0 = -n0;
      n1 = -n1;
      n2 = -n2;
    }
    const float len = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(n0, n0), __fmul_rn(n1, n1)), __fmul_rn(n2, n2)));
    if (len != 0.0f) {
      n0 = __fdiv_rn(n0, len);
      n1 = __fdiv_rn(n1, len);
      n2 = __fdiv_rn(n2, len);
    }
  }
  out4[orig] = make_float4(n0, n1, n2, curv);
}

// Radius covariances, COV_GRID_WGSL semantics: eigenvalues (eps, 1, 1) with the
// smallest mapped to eps; eps*I below 3 neighbours. One thread per point, the
// neighbourhood read from device memory as nine row ranges (the form without
// LDS tiles).
__global__ void __launch_bounds__(kBlock) k_radius_cov(GridDesc g, const uint32_t* __restrict__ cs,
                                                       const float4* __restrict__ gp, float r2,
                                                       float eps, float4* __restrict__ nrm_orig,
                                                       float* __restrict__ cov6,
                                                       float* __restrict__ normal3) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.npoints) return;
  const float4 p = gp[i];
  float a[3][3], v[3][3];
  const float count = radius_eigen_rows(g, cs, gp, p, r2, a, v);
  radius_cov_out(p, count, a, v, eps, nrm_orig, cov6, normal3);
}

// GpuNormalEstimator with a search radius (features/src/normal_gpu.rs:33-82):
// NORMALS_GRID_WGSL outputs (normals_grid.rs:160-176: unit normal of the
// smallest eigenvalue, curvature = lambda_min / trace; (0,0,1,0) below 3
// neighbours), then, when a viewpoint is given, orient_normal_towards_viewpoint
// (features/src/normal.rs:406-421). out4 in original point order.
__global__ void __launch_bounds__(kBlock) k_grid_normals(GridDesc g, const uint32_t* __restrict__ cs,
                                                         const float4* __restrict__ gp, float r2,
                                                         int has_vp, float vx, float vy, float vz,
                                                         float4* __restrict__ out4) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.npoints) return;
  const float4 p = gp[i];
  float a[3][3], v[3][3];
  const float count = radius_eigen_rows(g, cs, gp, p, r2, a, v);
  grid_normal_out(p, count, a, v, has_vp, vx, vy, vz, out4);
}

// ---- LDS-staged cell tiles ---------------------------------------------------
//
// A tile is a run of occupied cells of one grid row (at most kTileCells cells,
// about kTileTarget points: whole cells, so a cell holding more points is a
// tile of its own). Its points' 27-cell neighbourhoods all lie in the 3 x 3
// rows around it, cells x0-1 .. x0+len, which in grid order are nine
// contiguous point ranges; a workgroup stages them (and their cell starts) in
// LDS once, and every point of the tile then reads its neighbours from LDS in
// exactly the order of the per-point form, so the f32 sums and the outputs are
// bit-identical to it. A tile whose neighbourhood exceeds kHaloCap points is
// read from device memory instead (same arithmetic).
//
// Both forms read a point's neighbourhood as nine row ranges, four points per
// round. Measured on 898k points (radius 0.4 m): 0.46 ms against 0.19 ms for
// the per-point form, which reads the same neighbourhoods at L2 speed with far
// more waves in flight (cell by cell, the two took 0.63 and 0.29 ms); tiles of
// 128 or 256 threads, or larger halos, are slower still (the LDS per workgroup
// caps the waves per CU, and many rows crossing a wall make tiles of a few
// points). The per-point form is the default ("radius_tiles" 0); this one is
// kept selectable and tested equal.

#ifndef SR_TILE_THREADS
#define SR_TILE_THREADS 64
#endif
#ifndef SR_HALO_CAP
#define SR_HALO_CAP 1024
#endif
constexpr int kTileCells = 32;
constexpr int kTileThreads = SR_TILE_THREADS;
constexpr uint32_t kTileTarget = uint32_t(SR_TILE_THREADS);
constexpr uint32_t kHaloCap = SR_HALO_CAP;  // float4 points staged in LDS

struct RadiusTile {
  uint32_t row;  // cz * dy + cy
  uint32_t x0;   // first cell (x)
  uint32_t len;  // cells
  uint32_t pad_;
};

// One wave per grid row: the row's tiles, appended to `tiles`. The wave loads
// the point counts of 64 cells at a time (a lane a cell) and walks them in
// order, every lane keeping the same tile state: from the first occupied
// cell, cells (empty ones included) join the open tile until it spans
// kTileCells cells or the next cell would take an occupied tile past
// kTileTarget points; the tile then closes and the walk goes on from that
// cell. The walk runs twice: once to count the row's tiles, whose slots one
// atomic then reserves, once to write them. (A thread walking its row with two
// dependent loads a cell, and an atomic on the one counter per tile, took
// 0.70 ms for the 898k-point block at r = 0.4 m; this form 0.16 ms.)
__global__ void k_radius_tiles(GridDesc g, const uint32_t* __restrict__ cs,
                               RadiusTile* __restrict__ tiles, uint32_t* __restrict__ count) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t row = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;  // the same on a wave's lanes
  if (row >= uint32_t(g.dy) * uint32_t(g.dz)) return;
  const uint32_t base = row * uint32_t(g.dx), dx = uint32_t(g.dx);
  if (cs[base] == cs[base + dx]) return;
  auto walk = [&](auto&& emit) {
    bool open = false;
    uint32_t x0 = 0, pts = 0;
    for (uint32_t c0 = 0; c0 < dx; c0 += kWave) {
      const uint32_t x = c0 + lane;
      const uint32_t cnt = x < dx ? cs[base + x + 1] - cs[base + x] : 0u;
      const uint32_t m = min(dx - c0, uint32_t(kWave));
      for (uint32_t k = 0; k < m; ++k) {
        // (k is the same on every lane: a lane read, not a shuffle)
        const uint32_t xc = c0 + k, c = uint32_t(__builtin_amdgcn_readlane(int(cnt), int(k)));
        if (open && (xc - x0 >= uint32_t(kTileCells) || (pts > 0 && pts + c > kTileTarget))) {
          emit(x0, xc - x0);
          open = false;
        }
        if (!open) {
          if (c == 0) continue;
          open = true;
          x0 = xc;
          pts = 0;
        }
        pts += c;
      }
    }
    if (open) emit(x0, dx - x0);
  };
  uint32_t n = 0;
  walk([&](uint32_t, uint32_t) { ++n; });
  uint32_t slot = 0;
  if (lane == 0) slot = atomicAdd(count, n);
  slot = __shfl(slot, 0);
  walk([&](uint32_t x0, uint32_t len) {
    if (lane == 0) tiles[slot] = RadiusTile{row, x0, len, 0u};
    ++slot;
  });
}

// kMode 0: radius covariances; 1: grid normals. The grid has a fixed number of
// workgroups, each taking tiles in turn.
template <int kMode>
__global__ void __launch_bounds__(kTileThreads) k_radius_tiled(
    GridDesc g, const uint32_t* __restrict__ cs, const float4* __restrict__ gp,
    const RadiusTile* __restrict__ tiles, const uint32_t* __restrict__ ntiles, float r2, float eps,
    float4* __restrict__ nrm_orig, float* __restrict__ cov6, float* __restrict__ normal3,
    int has_vp, float vx, float vy, float vz, float4* __restrict__ out4) {
  __shared__ float4 halo[kHaloCap];
  __shared__ uint32_t cst[9][kTileCells + 3];  // per halo row: LDS offset of cells xa .. xb+1
  __shared__ uint32_t rlo[9], rcnt[9], roff[10];
  const uint32_t nt = *ntiles;
  for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
    const RadiusTile T = tiles[t];
    const int cy = int(T.row % uint32_t(g.dy)), cz = int(T.row / uint32_t(g.dy));
    const int xa = max(int(T.x0) - 1, 0), xb = min(int(T.x0 + T.len), g.dx - 1);
    const int ncell = xb - xa + 1;
    if (threadIdx.x < 9) {
      const int rr = threadIdx.x, ny = cy + rr % 3 - 1, nz = cz + rr / 3 - 1;
      uint32_t lo = 0, cnt = 0;
      if (ny >= 0 && ny < g.dy && nz >= 0 && nz < g.dz) {
        const uint32_t rb = (uint32_t(nz) * uint32_t(g.dy) + uint32_t(ny)) * uint32_t(g.dx);
        lo = cs[rb + uint32_t(xa)];
        cnt = cs[rb + uint32_t(xb) + 1] - lo;
      }
      rlo[rr] = lo;
      rcnt[rr] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t o = 0;
      for (int rr = 0; rr < 9; ++rr) {
        roff[rr] = o;
        o += rcnt[rr];
      }
      roff[9] = o;
    }
    __syncthreads();
    const bool staged = roff[9] <= kHaloCap;
    if (staged) {
      // one flat copy over the nine ranges, so all their loads are in flight together
      for (uint32_t k = threadIdx.x; k < roff[9]; k += kTileThreads) {
        int rr = 0;
#pragma unroll
        for (int q = 1; q < 9; ++q) rr += k >= roff[q] ? 1 : 0;
        halo[k] = gp[rlo[rr] + (k - roff[rr])];
      }
      for (int k = threadIdx.x; k < 9 * (ncell + 1); k += kTileThreads) {
        const int rr = k / (ncell + 1), j = k % (ncell + 1);
        const int ny = cy + rr % 3 - 1, nz = cz + rr / 3 - 1;
        uint32_t v = roff[rr];
        if (ny >= 0 && ny < g.dy && nz >= 0 && nz < g.dz) {
          const uint32_t rb = (uint32_t(nz) * uint32_t(g.dy) + uint32_t(ny)) * uint32_t(g.dx);
          v = roff[rr] + (cs[rb + uint32_t(xa + j)] - rlo[rr]);
        }
        cst[rr][j] = v;
      }
    }
    __syncthreads();
    const uint32_t rb0 = T.row * uint32_t(g.dx);
    const uint32_t p0 = cs[rb0 + T.x0], p1 = cs[rb0 + T.x0 + T.len];
    for (uint32_t i = p0 + threadIdx.x; i < p1; i += kTileThreads) {
      const float4 p = gp[i];
      float a[3][3], v[3][3];
      float count;
      if (staged) {
        // the point's nine row ranges in the halo (its rows lie within the
        // tile's 3 x 3 rows, its cells within xa .. xb)
        int py_, pz_;
        uint32_t x0, x1;
        radius_rows(g, p, py_, pz_, x0, x1);
        uint32_t lo[9], hi[9];
#pragma unroll
        for (int r = 0; r < 9; ++r) {
          const int rr = (pz_ + r / 3 - 1 - cz + 1) * 3 + (py_ + r % 3 - 1 - cy + 1);
          lo[r] = cst[rr][x0 - uint32_t(xa)];
          hi[r] = cst[rr][x1 - uint32_t(xa)];
        }
        count = radius_eigen_ranges(p, r2, lo, hi, [&](uint32_t s) { return halo[s]; }, a, v);
      } else {
        count = radius_eigen(g, cs, gp, p, r2, a, v);
      }
      if constexpr (kMode == 0) radius_cov_out(p, count, a, v, eps, nrm_orig, cov6, normal3);
      else grid_normal_out(p, count, a, v, has_vp, vx, vy, vz, out4);
    }
    __syncthreads();  // the LDS is reused by the next tile
  }
}

// Tiles of a radius grid (device list and count).
struct TileSet {
  DevBuf<RadiusTile> tiles;
  DevBuf<uint32_t> count;
  uint32_t grid = 0;  // workgroups of the tiled kernels
};

static void make_tiles(const Grid& rg, TileSet& ts, hipStream_t s) {
  const GridDesc& d = rg.desc;
  const uint64_t rows = uint64_t(d.dy) * uint64_t(d.dz);
  // a tile holds at least one point, and a row's tiles are at most its
  // occupied cells: min(points, cells) bounds the count
  const uint64_t cap = std::max<uint64_t>(std::min<uint64_t>(d.npoints, d.ncells), 1);
  ts.tiles.resize(cap);
  ts.count.resize(1);
  fill_u32(ts.count.p, 0u, 1, s);
  hipLaunchKernelGGL(k_radius_tiles, dim3(uint32_t((rows * kWave + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     s, d, rg.cell_start.p, ts.tiles.p, ts.count.p);
  // enough workgroups for several per CU; each takes tiles in turn
  ts.grid = uint32_t(std::min<uint64_t>(std::max<uint64_t>((d.npoints + kTileTarget - 1) / kTileTarget, 1), 16384));
}

void radius_normals(const Grid& rg, float radius, float eps, float4* nrm_orig, float* cov6,
                    float* normal3, hipStream_t s) {
  const uint32_t n = rg.desc.npoints;
  if (n == 0) return;
  if (options().radius_tiles.load()) {
    ProfScope prof("radius_cov", s);  // the tile list included
    TileSet ts;
    make_tiles(rg, ts, s);
    hipLaunchKernelGGL((k_radius_tiled<0>), dim3(ts.grid), dim3(kTileThreads), 0, s, rg.desc,
                       rg.cell_start.p, rg.pts.p, ts.tiles.p, ts.count.p, radius * radius, eps, nrm_orig,
                       cov6, normal3, 0, 0.0f, 0.0f, 0.0f, nullptr);
    return;
  }
  ProfScope prof("radius_cov", s);
  hipLaunchKernelGGL(k_radius_cov, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, rg.desc,
                     rg.cell_start.p, rg.pts.p, radius * radius, eps, nrm_orig, cov6, normal3);
}

void grid_normals(const Grid& rg, float radius, const float* viewpoint, float4* out_orig,
                  hipStream_t s) {
  const uint32_t n = rg.desc.npoints;
  if (n == 0) return;
  const int has_vp = viewpoint ? 1 : 0;
  const float vx = viewpoint ? viewpoint[0] : 0.0f, vy = viewpoint ? viewpoint[1] : 0.0f,
              vz = viewpoint ? viewpoint[2] : 0.0f;
  if (options().radius_tiles.load()) {
    ProfScope prof("grid_normals", s);  // the tile list included
    TileSet ts;
    make_tiles(rg, ts, s);
    hipLaunchKernelGGL((k_radius_tiled<1>), dim3(ts.grid), dim3(kTileThreads), 0, s, rg.desc,
                       rg.cell_start.p, rg.pts.p, ts.tiles.p, ts.count.p, radius * radius, 0.0f, nullptr,
                       nullptr, nullptr, has_vp, vx, vy, vz, out_orig);
    return;
  }
  ProfScope prof("grid_normals", s);
  hipLaunchKernelGGL(k_grid_normals, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, rg.desc,
                     rg.cell_start.p, rg.pts.p, radius * radius, has_vp, vx, vy, vz, out_orig);
}

}  // namespace sr
