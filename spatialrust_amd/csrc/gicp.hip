// One GICP Gauss-Newton pass per launch, for any number of registrations.
//
// This is the GPU form of the iteration body of
// GicpRegistration::align_with_diagnostics
// (crates/spatialrust-registration/src/gicp.rs:135-209):
//   p' = R p + t (f32), j = nearest target point, gate d2 > dmax^2 (f32),
//   C = C_t[j] + R C_s[i] R^T, M = C^-1, e = p' - q_j, J = [-skew(p') | I],
//   H += J^T M J, g += J^T M e, count += 1;
//   then H x = -g, pose <- delta(x) o pose, convergence checks.
// Differences of form (results agree to the tolerance the tests state):
//  - C_t, C_s are stored as unit normals: C = a I - (1-eps) n n^T (a = 1, or
//    a = eps with n = 0 for the radius path's eps*I), so M has a closed form
//    (no 3x3 adjugate, which in f32 would lose most digits at cond(C) ~ 1e3);
//  - per-point terms in f32; the workgroup reduces its threads' sums in a fixed
//    order with Kahan-compensated chunks and a TwoSum combine, writing (hi, lo)
//    pairs; the finalize kernel adds the block partials in FP64 in a fixed order
//    — so a registration's result does not depend on what else shares the launch;
//  - the 6x6 solve is an FP64 LDL^T (the reference runs Gaussian elimination
//    with partial pivoting, linalg.rs:12-53; both stop on pivots below 1e-12);
//  - the pose is held in FP64 and updated with the reference's SO(3) x R^3
//    left update (delta_from_solution + Isometry3::compose, gicp.rs:314-324,
//    transform.rs:104-108);
//  - the fitness test of iteration k (a second nearest-neighbour pass in the
//    reference, gicp.rs:203-208, 330-357) is taken from pass k+1, which visits
//    the same transformed points against the same target.
// A pass is two launches — the point-parallel pass and a one-workgroup-per-
// registration finalize — and the host never waits between iterations. (A
// last-workgroup-finalizes scheme needs a device-scope fence in every
// workgroup, which on gfx950 writes back and invalidates the XCD's L2 and cost
// more than the extra launch.)

#include <cfloat>

#include "internal.hpp"

namespace sr {

__device__ __forceinline__ void q_to_mat64(const double q0[4], double R[3][3]) {
  double len = sqrt(q0[0] * q0[0] + q0[1] * q0[1] + q0[2] * q0[2] + q0[3] * q0[3]);
  double x = 0, y = 0, z = 0, w = 1;
  if (len != 0.0) {
    x = q0[0] / len;
    y = q0[1] / len;
    z = q0[2] / len;
    w = q0[3] / len;
  }
  double xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
  double wx = w * x, wy = w * y, wz = w * z;
  // Quat::to_mat3 (crates/spatialrust-math/src/quat.rs:53-71) after its transpose.
  R[0][0] = 1.0 - 2.0 * (yy + zz);
  R[0][1] = 2.0 * (xy - wz);
  R[0][2] = 2.0 * (xz + wy);
  R[1][0] = 2.0 * (xy + wz);
  R[1][1] = 1.0 - 2.0 * (xx + zz);
  R[1][2] = 2.0 * (yz - wx);
  R[2][0] = 2.0 * (xz - wy);
  R[2][1] = 2.0 * (yz + wx);
  R[2][2] = 1.0 - 2.0 * (xx + yy);
}

__device__ __forceinline__ void refresh_f32_pose(InstState* st) {
  double R[3][3];
  q_to_mat64(st->q, R);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      const float hi = float(R[a][b]);
      st->rf[3 * a + b] = hi;
      st->rlo[3 * a + b] = float(R[a][b] - double(hi));
    }
  for (int a = 0; a < 3; ++a) {
    const float hi = float(st->t[a]);
    st->tf[a] = hi;
    st->tlo[a] = float(st->t[a] - double(hi));
  }
}

__global__ void k_init_states(InstState* states, const double* __restrict__ init7, uint32_t count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  InstState s;
  for (int a = 0; a < 4; ++a) s.q[a] = init7[7 * i + a];
  for (int a = 0; a < 3; ++a) s.t[a] = init7[7 * i + 4 + a];
  s.fitness = DBL_MAX;
  s.last_update = 0.0;
  s.count = 0;
  s.phase = 0;
  s.iter = 0;
  s.status = 0;
  s.converged = 0;
  s.counter = 0;
  s.pad_ = 0;
  refresh_f32_pose(&s);
  states[i] = s;
}

void init_states(InstState* states, const double* init7_dev, uint32_t count, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_init_states, dim3((count + 127) / 128), dim3(128), 0, s, states, init7_dev,
                     count);
}

// FP64 LDL^T solve of H x = -g (H symmetric, packed upper). false if a pivot
// falls below 1e-12 — the singularity threshold of solve_linear_system.
// The factor lives in LDS scratch `w` (>= 48 doubles) rather than registers:
// this runs on one thread of one workgroup per pass, and keeping 6x6 FP64
// arrays in VGPRs would cap the occupancy of the whole point-parallel kernel.
__device__ bool ldlt_solve(const double* Hp, const double* g, double* x, double* w) {
  double* L = w;       // 36, row-major, unit lower
  double* D = w + 36;  // 6
  double* y = w + 42;  // 6
  for (int j = 0; j < 6; ++j) {
    double d = Hp[ut(j, j)];
    for (int k = 0; k < j; ++k) d -= L[6 * j + k] * L[6 * j + k] * D[k];
    if (!(d >= 1e-12)) return false;
    D[j] = d;
    L[6 * j + j] = 1.0;
    for (int i = j + 1; i < 6; ++i) {
      double v = Hp[ut(j, i)];
      for (int k = 0; k < j; ++k) v -= L[6 * i + k] * L[6 * j + k] * D[k];
      L[6 * i + j] = v / d;
    }
  }
  for (int i = 0; i < 6; ++i) {
    double v = -g[i];
    for (int k = 0; k < i; ++k) v -= L[6 * i + k] * y[k];
    y[i] = v;
  }
  for (int i = 5; i >= 0; --i) {
    double v = y[i] / D[i];
    for (int k = i + 1; k < 6; ++k) v -= L[6 * k + i] * x[k];
    x[i] = v;
  }
  return true;
}

// Gauss-Newton bookkeeping of one registration after a pass (one thread).
__device__ __forceinline__ void finalize(InstState* st, const InstConfig& cfg, const double* Hp,
                                      const double* g, double fit_sum, unsigned long long cnt,
                                      double* w) {
  const double F = cnt > 0 ? fit_sum / double(cnt) : DBL_MAX;
  st->count = cnt;
  if (st->phase == 1) {  // final fitness pass after transformation_epsilon convergence
    st->fitness = F;
    st->phase = 2;
    return;
  }
  if (st->iter >= 1 && F < cfg.fit_eps) {  // gicp.rs:203-208
    st->fitness = F;
    st->converged = 1;
    st->phase = 2;
    return;
  }
  if (st->iter >= cfg.max_iter) {  // loop exhausted (gicp.rs:135, 211-220)
    st->fitness = F;
    st->phase = 2;
    return;
  }
  st->iter += 1;
  if (cnt < cfg.min_corr) {  // gicp.rs:181-186
    st->status = 1;
    st->phase = 2;
    return;
  }
  double* x = w + 48;
  if (!ldlt_solve(Hp, g, x, w)) {  // gicp.rs:190-194
    st->status = 2;
    st->phase = 2;
    return;
  }
  // delta_from_solution (gicp.rs:314-324) and Isometry3::compose (transform.rs:104-108), FP64.
  const double angle = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  double dq[4] = {0.0, 0.0, 0.0, 1.0};
  if (angle > 1e-12) {
    double ax = x[0] / angle, ay = x[1] / angle, az = x[2] / angle;
    double al = sqrt(ax * ax + ay * ay + az * az);
    ax /= al;
    ay /= al;
    az /= al;
    double sh, ch;
    sincos(0.5 * angle, &sh, &ch);
    dq[0] = ax * sh;
    dq[1] = ay * sh;
    dq[2] = az * sh;
    dq[3] = ch;
    double l = sqrt(dq[0] * dq[0] + dq[1] * dq[1] + dq[2] * dq[2] + dq[3] * dq[3]);
    for (int a = 0; a < 4; ++a) dq[a] /= l;
  }
  const double* q = st->q;
  double nq[4] = {dq[3] * q[0] + dq[0] * q[3] + dq[1] * q[2] - dq[2] * q[1],
                  dq[3] * q[1] - dq[0] * q[2] + dq[1] * q[3] + dq[2] * q[0],
                  dq[3] * q[2] + dq[0] * q[1] - dq[1] * q[0] + dq[2] * q[3],
                  dq[3] * q[3] - dq[0] * q[0] - dq[1] * q[1] - dq[2] * q[2]};
  double l = sqrt(nq[0] * nq[0] + nq[1] * nq[1] + nq[2] * nq[2] + nq[3] * nq[3]);
  double Rd[3][3];
  q_to_mat64(dq, Rd);
  double nt[3];
  for (int a = 0; a < 3; ++a)
    nt[a] = Rd[a][0] * st->t[0] + Rd[a][1] * st->t[1] + Rd[a][2] * st->t[2] + x[3 + a];
  for (int a = 0; a < 4; ++a) st->q[a] = nq[a] / l;
  for (int a = 0; a < 3; ++a) st->t[a] = nt[a];
  refresh_f32_pose(st);
  double upd = 0.0;  // update_magnitude (gicp.rs:326-328)
  for (int a = 0; a < 6; ++a) upd += x[a] * x[a];
  upd = sqrt(upd);
  st->last_update = upd;
  if (upd < cfg.trans_eps) {  // gicp.rs:199-202
    st->converged = 1;
    st->phase = 1;
  }
}

// Nearest target point of p within dmax (exact; ties to the lower index).
// Returns the grid position or kNone. `prev` (grid position or kNone) is the
// previous pass's correspondence of this source point: a real candidate, so it
// seeds the best and the search radius without changing the answer.
template <bool kStats>
__device__ __forceinline__ uint32_t nearest_within(const GridDesc& g, const uint32_t* __restrict__ cs,
                                                   const uint8_t* __restrict__ empty,
                                                   const float4* __restrict__ tp, float px, float py,
                                                   float pz, float dmax, float dmax2, uint32_t prev,
                                                   float& best_d, SearchStats* stats) {
  best_d = INFINITY;
  uint32_t best_pos = kNone;
  uint32_t n_rings = 0, n_cells = 0, n_pts = 0;
  // Quick reject: query farther than dmax from the grid's box.
  {
    float bx = fmaxf(fmaxf(g.ox - px, px - (g.ox + g.h * g.dx)), 0.0f);
    float by = fmaxf(fmaxf(g.oy - py, py - (g.oy + g.h * g.dy)), 0.0f);
    float bz = fmaxf(fmaxf(g.oz - pz, pz - (g.oz + g.h * g.dz)), 0.0f);
    float lb = bx * bx + by * by + bz * bz;
    if (lb > dmax2 + 2.0f * dmax * g.margin + g.margin * g.margin) {
      if constexpr (kStats) stats_add(stats, 0, 0, 0, kRejected);
      return kNone;
    }
  }
  const int hx = cell_coord_free(px, g.ox, g.inv_h), hy = cell_coord_free(py, g.oy, g.inv_h),
            hz = cell_coord_free(pz, g.oz, g.inv_h);
  const bool inside = hx >= 0 && hx < g.dx && hy >= 0 && hy < g.dy && hz >= 0 && hz < g.dz;
  int e = 0;
  if (empty && inside) {
    // Rings closer than the empty radius hold no point, so every point lies
    // outside the box of rings 0..e-1: its clearance bounds the distance.
    e = empty[(uint32_t(hz) * uint32_t(g.dy) + uint32_t(hy)) * uint32_t(g.dx) + uint32_t(hx)];
    if (e > 0 && ring_clearance(g, px, py, pz, hx, hy, hz, e - 1) > dmax) {
      if constexpr (kStats) stats_add(stats, 0, 0, 0, kRejected);
      return kNone;
    }
  }
  // thr = min(best so far, dmax^2): cells and rows beyond it are never touched.
  float thr = dmax2;
  // (d2, original index) packed as one u64 key: d2 >= 0, so its bits order like
  // the float, and the index breaks ties. The initial key (dmax2, ~0) admits
  // exactly d2 <= dmax2 (NaN distances compare above it).
  uint64_t best_key = (uint64_t(__float_as_uint(dmax2)) << 32) | 0xFFFFFFFFull;
  auto consider = [&](uint32_t s, const float4& q) {
    const float d2 = dist2(px, py, pz, q.x, q.y, q.z);
    const uint64_t key = (uint64_t(__float_as_uint(d2)) << 32) | __float_as_uint(q.w);
    if (key < best_key) {
      best_key = key;
      best_pos = s;
      thr = d2;
    }
  };
  if (prev != kNone) consider(prev, tp[prev]);
  if (inside && e <= 1) {
    // Rings 0 and 1 as nine contiguous point ranges, cut to the sphere of the
    // seeded radius; most queries of a converging registration end here.
    if constexpr (kStats) ++n_rings;
    scan_block27_within(g, cs, tp, px, py, pz, hx, hy, hz, thr, [&](uint32_t s, const float4& q) {
      if constexpr (kStats) ++n_pts;
      consider(s, q);
    });
    const float cl = ring_clearance(g, px, py, pz, hx, hy, hz, 1);
    if (cl > 0.0f && thr < cl * cl) {
      if constexpr (kStats) stats_add(stats, n_rings, n_cells, n_pts, kBlockDone);
      if (best_pos != kNone) best_d = __uint_as_float(uint32_t(best_key >> 32));
      return best_pos;
    }
  }
  // Not certified by the block: every row of the sphere of the current radius,
  // each once (the block's cells excluded when it was scanned).
  if constexpr (kStats) ++n_rings;
  scan_ball(g, cs, tp, px, py, pz, hx, hy, hz, thr, inside && e <= 1, n_cells,
            [&](uint32_t s, const float4& q) {
              if constexpr (kStats) ++n_pts;
              consider(s, q);
            });
  if constexpr (kStats)
    stats_add(stats, n_rings, n_cells, n_pts, best_pos == kNone ? kRingNone : kRingFound);
  if (best_pos != kNone) best_d = __uint_as_float(uint32_t(best_key >> 32));
  return best_pos;
}

struct PointTerms {
  float v[kTerms];
};

// Closed-form M = (C_t + R C_s R^T)^-1 with C = a I - beta n n^T.
__device__ __forceinline__ void combined_inverse(float ux, float uy, float uz, float at, float vx,
                                                 float vy, float vz, float as, float beta,
                                                 float eps, float M[6]) {
  // M packed: m00 m01 m02 m11 m12 m22
  if (at == 1.0f && as == 1.0f) {
    // Both planar: C = 2I - beta (u u^T + v v^T), |u| = |v| = 1. With s = u+v,
    // d = u-v (orthogonal), M = I/2 + ks s s^T + kd d d^T where
    // ks = beta / (8 eps + 2 beta |d|^2), kd = beta / (8 - 2 beta |d|^2).
    const float sx = ux + vx, sy = uy + vy, sz = uz + vz;
    const float dx = ux - vx, dy = uy - vy, dz = uz - vz;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float ks = beta / (8.0f * eps + 2.0f * beta * d2);
    const float kd = beta / (8.0f - 2.0f * beta * d2);
    M[0] = 0.5f + ks * sx * sx + kd * dx * dx;
    M[1] = ks * sx * sy + kd * dx * dy;
    M[2] = ks * sx * sz + kd * dx * dz;
    M[3] = 0.5f + ks * sy * sy + kd * dy * dy;
    M[4] = ks * sy * sz + kd * dy * dz;
    M[5] = 0.5f + ks * sz * sz + kd * dz * dz;
    return;
  }
  // At least one side is eps*I (n = 0): C = alpha I - beta w w^T with w the
  // other side's normal (or 0); Sherman-Morrison.
  const float alpha = at + as;
  float wx = 0, wy = 0, wz = 0, bw = 0.0f;
  if (at == 1.0f) {
    wx = ux;
    wy = uy;
    wz = uz;
    bw = beta;
  } else if (as == 1.0f) {
    wx = vx;
    wy = vy;
    wz = vz;
    bw = beta;
  }
  const float ia = 1.0f / alpha;
  const float k = bw / (alpha * (alpha - bw));
  M[0] = ia + k * wx * wx;
  M[1] = k * wx * wy;
  M[2] = k * wx * wz;
  M[3] = ia + k * wy * wy;
  M[4] = k * wy * wz;
  M[5] = ia + k * wz * wz;
}

template <bool kStats>
__global__ void __launch_bounds__(kBlock) k_gicp_pass(const IterInstance* __restrict__ insts,
                                                      const uint32_t* __restrict__ block_inst,
                                                      SearchStats* __restrict__ stats) {
  // XCD-aware block order: the hardware deals consecutive workgroup ids round
  // robin over the 8 XCDs; give each XCD a contiguous range of logical blocks so
  // neighbouring source tiles (which touch the same target cells) share an L2.
  const uint32_t G = gridDim.x, hw = blockIdx.x;
  const uint32_t xcd = hw & 7u, q8 = G >> 3, rem = G & 7u;
  const uint32_t b = xcd * q8 + min(xcd, rem) + (hw >> 3);

  const uint32_t inst_id = block_inst[b];
  const IterInstance& I = insts[inst_id];
  InstState* st = I.state;
  if (st->phase == 2) return;
  const uint32_t lb = b - I.block_begin;
  const GridDesc g = I.grid;
  const InstConfig cfg = I.cfg;
  const float beta = cfg.one_minus_eps, eps = 1.0f - cfg.one_minus_eps;
  float R[9], T[3], Rl[9], Tl[3];
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    R[a] = st->rf[a];
    Rl[a] = st->rlo[a];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    T[a] = st->tf[a];
    Tl[a] = st->tlo[a];
  }

  // Plain f32 per-thread sums (a thread holds only ppt points); compensation
  // happens in the block reduction below.
  float acc[kTerms];
#pragma unroll
  for (int t = 0; t < kTerms; ++t) acc[t] = 0.0f;
  uint32_t count = 0;

  const uint32_t ppt = I.ppt;
  for (uint32_t k = 0; k < ppt; ++k) {
    const uint32_t i = (lb * ppt + k) * kBlock + threadIdx.x;
    if (i >= I.n_src) break;
    const float4 p = I.src[i];
    // p' = R p + t (Isometry3::transform_point, transform.rs:126-128) with the
    // FP64 pose split into f32 high and low parts: the low part removes the
    // systematic error of an f32-rounded pose, which would otherwise floor the
    // Gauss-Newton update at ~1e-7 (the reference's own f32 pose has that floor).
    const float lx = fmaf(Rl[0], p.x, fmaf(Rl[1], p.y, fmaf(Rl[2], p.z, Tl[0])));
    const float ly = fmaf(Rl[3], p.x, fmaf(Rl[4], p.y, fmaf(Rl[5], p.z, Tl[1])));
    const float lz = fmaf(Rl[6], p.x, fmaf(Rl[7], p.y, fmaf(Rl[8], p.z, Tl[2])));
    const float px = fmaf(R[0], p.x, fmaf(R[1], p.y, fmaf(R[2], p.z, T[0] + lx)));
    const float py = fmaf(R[3], p.x, fmaf(R[4], p.y, fmaf(R[5], p.z, T[1] + ly)));
    const float pz = fmaf(R[6], p.x, fmaf(R[7], p.y, fmaf(R[8], p.z, T[2] + lz)));
    float d2;
    const uint32_t j = nearest_within<kStats>(g, I.cell_start, I.empty_radius, I.tgt, px, py, pz, cfg.dmax,
                                      cfg.dmax2, I.match[i], d2, stats);
    I.match[i] = j;
    if (j == kNone) continue;
    const float4 q = I.tgt[j];
    const float4 nt = I.tgt_n[j];
    const float4 ns = I.src_n[i];
    const float vx = R[0] * ns.x + R[1] * ns.y + R[2] * ns.z;
    const float vy = R[3] * ns.x + R[4] * ns.y + R[5] * ns.z;
    const float vz = R[6] * ns.x + R[7] * ns.y + R[8] * ns.z;
    float M[6];
    combined_inverse(nt.x, nt.y, nt.z, nt.w, vx, vy, vz, ns.w, beta, eps, M);
    const float ex = px - q.x, ey = py - q.y, ez = pz - q.z;
    // w = M e
    const float wx = M[0] * ex + M[1] * ey + M[2] * ez;
    const float wy = M[1] * ex + M[3] * ey + M[4] * ez;
    const float wz = M[2] * ex + M[4] * ey + M[5] * ez;
    // A = skew(p') M ; H_rt = A ; H_rr = -A skew(p') ; H_tt = M
    const float m[3][3] = {{M[0], M[1], M[2]}, {M[1], M[3], M[4]}, {M[2], M[4], M[5]}};
    float A[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      A[0][c] = -pz * m[1][c] + py * m[2][c];
      A[1][c] = pz * m[0][c] - px * m[2][c];
      A[2][c] = -py * m[0][c] + px * m[1][c];
    }
    float Hrr[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      Hrr[a][0] = -(A[a][1] * pz - A[a][2] * py);
      Hrr[a][1] = -(-A[a][0] * pz + A[a][2] * px);
      Hrr[a][2] = -(A[a][0] * py - A[a][1] * px);
    }
    // H upper triangle, row-major over the 6x6.
    acc[ut(0, 0)] += (Hrr[0][0]);
    acc[ut(0, 1)] += (Hrr[0][1]);
    acc[ut(0, 2)] += (Hrr[0][2]);
    acc[ut(1, 1)] += (Hrr[1][1]);
    acc[ut(1, 2)] += (Hrr[1][2]);
    acc[ut(2, 2)] += (Hrr[2][2]);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[ut(a, 3 + c)] += (A[a][c]);
    acc[ut(3, 3)] += (M[0]);
    acc[ut(3, 4)] += (M[1]);
    acc[ut(3, 5)] += (M[2]);
    acc[ut(4, 4)] += (M[3]);
    acc[ut(4, 5)] += (M[4]);
    acc[ut(5, 5)] += (M[5]);
    // g = [p' x w ; w]
    acc[21] += (py * wz - pz * wy);
    acc[22] += (pz * wx - px * wz);
    acc[23] += (px * wy - py * wx);
    acc[24] += (wx);
    acc[25] += (wy);
    acc[26] += (wz);
    acc[27] += (d2);
    ++count;
  }

  // ---- block reduction (fixed order, compensated) -------------------------
  // Rows padded by 8 floats and chunks read interleaved (element e*8 + ch), so
  // the 64 lanes of a wave spread over the LDS banks.
  constexpr int kRow = kBlock + 8;
  __shared__ float red[kTerms][kRow];
  __shared__ float part_hi[kTerms][8], part_lo[kTerms][8];
  __shared__ unsigned long long wcount[kBlock / kWave];
#pragma unroll
  for (int t = 0; t < kTerms; ++t) red[t][threadIdx.x] = acc[t];
  unsigned long long c64 = count;
  for (int off = kWave / 2; off > 0; off >>= 1) c64 += __shfl_xor(c64, off);
  if ((threadIdx.x & (kWave - 1)) == 0) wcount[threadIdx.x / kWave] = c64;
  __syncthreads();
  if (threadIdx.x < kTerms * 8) {
    const int t = threadIdx.x >> 3, ch = threadIdx.x & 7;
    Kahan k;
    k.init();
    for (int e = 0; e < kBlock / 8; ++e) k.add(red[t][e * 8 + ch]);
    part_hi[t][ch] = k.s;
    part_lo[t][ch] = -k.c;
  }
  __syncthreads();
  // Block partial: 28 (hi, lo) pairs + u64 count. The finalize kernel of this
  // pass reads them after the kernel boundary, so no device-scope fence (an L2
  // write-back + invalidate on every XCD) is needed here.
  float* out = I.partials + size_t(lb) * 64;
  if (threadIdx.x < kTerms) {
    const int t = threadIdx.x;
    float hi = 0.0f, lo = 0.0f;
    for (int ch = 0; ch < 8; ++ch) {
      float s, e;
      two_sum(hi, part_hi[t][ch], s, e);
      hi = s;
      lo += e + part_lo[t][ch];
    }
    float s, e;
    two_sum(hi, lo, s, e);
    out[2 * t] = s;
    out[2 * t + 1] = e;
  }
  if (threadIdx.x == kTerms) {
    unsigned long long tot = 0;
    for (int w = 0; w < kBlock / kWave; ++w) tot += wcount[w];
    reinterpret_cast<unsigned long long*>(out)[30] = tot;  // floats 60..61
  }
}

// One workgroup per registration: FP64 sum of its block partials, then the
// Gauss-Newton bookkeeping, solve and pose update on one thread. Every thread
// takes blocks tid, tid+256, ... (many loads in flight), then a fixed
// xor-butterfly per wave and a fixed wave order: the same order on every run,
// whatever the launch mix.
__global__ void __launch_bounds__(kBlock) k_gicp_finalize(const IterInstance* __restrict__ insts,
                                                          uint32_t count) {
  const uint32_t inst_id = blockIdx.x;
  if (inst_id >= count) return;
  const IterInstance& I = insts[inst_id];
  InstState* st = I.state;
  if (st->phase == 2) return;
  const InstConfig cfg = I.cfg;
  // The state is staged in LDS: the serial bookkeeping below reads and writes
  // it field by field, which from global memory would chain many round trips.
  static_assert(sizeof(InstState) % 8 == 0 && sizeof(InstState) / 8 <= kBlock, "state staging");
  constexpr int kStateWords = int(sizeof(InstState) / 8);
  __shared__ InstState sst;
  if (threadIdx.x < kStateWords)
    reinterpret_cast<unsigned long long*>(&sst)[threadIdx.x] =
        reinterpret_cast<const unsigned long long*>(st)[threadIdx.x];
  __shared__ double wsum[kBlock / kWave][kTerms];
  __shared__ unsigned long long wcnt[kBlock / kWave];
  __shared__ double dtot[kTerms];
  __shared__ double work[64];
  const uint32_t nb = I.nblocks;
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  // All 28 terms of a block partial are loaded together (14 float4 + the
  // count), so each thread has its whole share of loads in flight at once.
  double v[kTerms];
#pragma unroll
  for (int t = 0; t < kTerms; ++t) v[t] = 0.0;
  unsigned long long cs = 0;
  for (uint32_t bb = threadIdx.x; bb < nb; bb += kBlock) {
    const float4* pp = reinterpret_cast<const float4*>(I.partials + size_t(bb) * 64);
    float4 q[kTerms / 2];
#pragma unroll
    for (int j = 0; j < kTerms / 2; ++j) q[j] = pp[j];
    cs += reinterpret_cast<const unsigned long long*>(pp)[30];
#pragma unroll
    for (int j = 0; j < kTerms / 2; ++j) {
      v[2 * j] += double(q[j].x) + double(q[j].y);
      v[2 * j + 1] += double(q[j].z) + double(q[j].w);
    }
  }
#pragma unroll
  for (int t = 0; t < kTerms; ++t)
    for (int off = kWave / 2; off > 0; off >>= 1) v[t] += __shfl_xor(v[t], off);
  for (int off = kWave / 2; off > 0; off >>= 1) cs += __shfl_xor(cs, off);
  if (lane == 0) {
#pragma unroll
    for (int t = 0; t < kTerms; ++t) wsum[wid][t] = v[t];
    wcnt[wid] = cs;
  }
  __syncthreads();
  if (threadIdx.x < kTerms) {
    double sacc = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) sacc += wsum[w][threadIdx.x];
    dtot[threadIdx.x] = sacc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long cnt = 0;
    for (int w = 0; w < kBlock / kWave; ++w) cnt += wcnt[w];
    finalize(&sst, cfg, dtot, dtot + 21, dtot[27], cnt, work);
  }
  __syncthreads();
  if (threadIdx.x < kStateWords)
    reinterpret_cast<unsigned long long*>(st)[threadIdx.x] =
        reinterpret_cast<const unsigned long long*>(&sst)[threadIdx.x];
}

void gicp_pass(const IterInstance* insts, const uint32_t* block_inst, uint32_t total_blocks,
               uint32_t count, hipStream_t s, bool finalize) {
  if (total_blocks == 0) return;
  {
    ProfScope prof("gicp_pass", s);
    SearchStats* st = prof_stats("gicp_pass");
    if (st)
      hipLaunchKernelGGL(k_gicp_pass<true>, dim3(total_blocks), dim3(kBlock), 0, s, insts,
                         block_inst, st);
    else
      hipLaunchKernelGGL(k_gicp_pass<false>, dim3(total_blocks), dim3(kBlock), 0, s, insts,
                         block_inst, st);
  }
  if (!finalize) return;
  ProfScope prof("gicp_finalize", s);
  hipLaunchKernelGGL(k_gicp_finalize, dim3(count), dim3(kBlock), 0, s, insts, count);
}

}  // namespace sr
