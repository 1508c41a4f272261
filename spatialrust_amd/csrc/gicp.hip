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

// 1/x and 1/sqrt(x) on the finalize's serial FP64 path: the hardware estimate
// refined by two Newton steps (each at least doubles the correct bits, so the
// result is within an ulp or two), in place of the IEEE-exact sequences, whose
// operand scaling and fix-up steps the well-scaled operands here (pivots
// >= 1e-12, lengths near 1) do not need.
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ double rsq64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  double t = fma(-h * y, y, 0.5);  // y <- y (3 - x y^2) / 2
  y = fma(y, t, y);
  t = fma(-h * y, y, 0.5);
  return fma(y, t, y);
}

__device__ __forceinline__ void q_to_mat64(const double q0[4], double R[3][3]) {
  const double n2 =
      fma(q0[0], q0[0], fma(q0[1], q0[1], fma(q0[2], q0[2], q0[3] * q0[3])));
  double x = 0, y = 0, z = 0, w = 1;
  if (n2 != 0.0) {
    // one reciprocal square root, four products
    const double il = n2 >= 1e-280 ? rsq64(n2) : 1.0 / sqrt(n2);
    x = q0[0] * il;
    y = q0[1] * il;
    z = q0[2] * il;
    w = q0[3] * il;
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

__device__ __forceinline__ void init_state(InstState* states, const double* __restrict__ init7,
                                           uint32_t i) {
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
  s.move_rot = INFINITY;
  s.move_const = INFINITY;
  refresh_f32_pose(&s);
  states[i] = s;
}

__global__ void k_init_states(InstState* states, const double* __restrict__ init7, uint32_t count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) init_state(states, init7, i);
}

void init_states(InstState* states, const double* init7_dev, uint32_t count, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_init_states, dim3((count + 127) / 128), dim3(128), 0, s, states, init7_dev,
                     count);
}

// A batch's starting state in one launch: each registration's state from its
// initial pose, and each source point without a match (kNone), budget (0) or
// searched position (NaN).
__global__ void k_reset_batch(InstState* states, const double* __restrict__ init7, uint32_t count,
                              uint32_t* __restrict__ match, float* __restrict__ budget,
                              float4* __restrict__ seen, uint64_t n) {
  const uint64_t t0 = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t0 < count) init_state(states, init7, uint32_t(t0));
  const float nan = __uint_as_float(0xFFFFFFFFu);
  for (uint64_t i = t0; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    match[i] = kNone;
    budget[i] = 0.0f;
    seen[i] = make_float4(nan, nan, nan, nan);
  }
}

void reset_batch_state(InstState* states, const double* init7_dev, uint32_t count, uint32_t* match,
                       float* budget, float4* seen, uint64_t n, hipStream_t s) {
  const uint64_t want = (std::max<uint64_t>(n, count) + kBlock - 1) / kBlock;
  const uint32_t blocks = uint32_t(std::min<uint64_t>(std::max<uint64_t>(want, 1), 4096));
  hipLaunchKernelGGL(k_reset_batch, dim3(blocks), dim3(kBlock), 0, s, states, init7_dev, count, match,
                     budget, seen, n);
}

// FP64 LDL^T solve of H x = -g (H symmetric, packed upper). false if a pivot
// falls below 1e-12 — the singularity threshold of solve_linear_system.
// Every loop bound is a constant, so the factor stays in registers: this runs
// on one thread of the finalize kernel, whose registers cost nothing else.
__device__ __forceinline__ bool ldlt_solve(const double* Hp, const double* g, double* x) {
  // LD[j][k] = L[j][k] D[k]: each elimination term is one fused product
  double L[6][6], LD[6][6], iD[6], y[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double d = Hp[ut(j, j)];
#pragma unroll
    for (int k = 0; k < j; ++k) d = fma(-L[j][k], LD[j][k], d);
    if (!(d >= 1e-12)) return false;
    iD[j] = rcp64(d);  // one reciprocal per pivot; the column takes products
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      double v = Hp[ut(j, i)];
#pragma unroll
      for (int k = 0; k < j; ++k) v = fma(-L[i][k], LD[j][k], v);
      LD[i][j] = v;
      L[i][j] = v * iD[j];
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double v = -g[i];
#pragma unroll
    for (int k = 0; k < i; ++k) v = fma(-L[i][k], y[k], v);
    y[i] = v;
  }
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double v = y[i] * iD[i];
#pragma unroll
    for (int k = i + 1; k < 6; ++k) v = fma(-L[k][i], x[k], v);
    x[i] = v;
  }
  return true;
}

// Gauss-Newton bookkeeping of one registration after a pass (one thread).
__device__ __forceinline__ void finalize(InstState* st, const InstConfig& cfg, const double* Hp,
                                      const double* g, double fit_sum, unsigned long long cnt) {
  const double F = cnt > 0 ? fit_sum / double(cnt) : DBL_MAX;
  st->count = cnt;
  st->move_rot = 0.0f;  // unchanged pose unless updated below
  st->move_const = 0.0f;
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
  double x[6];
  if (!ldlt_solve(Hp, g, x)) {  // gicp.rs:190-194
    st->status = 2;
    st->phase = 2;
    return;
  }
  // delta_from_solution (gicp.rs:314-324) and Isometry3::compose (transform.rs:104-108), FP64.
  const double angle = sqrt(fma(x[0], x[0], fma(x[1], x[1], x[2] * x[2])));
  double dq[4] = {0.0, 0.0, 0.0, 1.0};
  if (angle > 1e-12) {
    const double ia = rcp64(angle);
    double ax = x[0] * ia, ay = x[1] * ia, az = x[2] * ia;
    const double ial = rsq64(fma(ax, ax, fma(ay, ay, az * az)));
    ax *= ial;
    ay *= ial;
    az *= ial;
    double sh, ch;
    const double h = 0.5 * angle;
    if (h < 0.0625) {
      // Taylor series to x^17 / x^16: below 1/16 the first omitted terms are
      // under 2^-100 relative, so this is sin / cos to FP64 rounding without
      // the general argument reduction (Horner form, one fused step a term).
      const double h2 = h * h;
      double ps = 1.0 / 355687428096000.0;
      ps = fma(h2, ps, -1.0 / 1307674368000.0);
      ps = fma(h2, ps, 1.0 / 6227020800.0);
      ps = fma(h2, ps, -1.0 / 39916800);
      ps = fma(h2, ps, 1.0 / 362880);
      ps = fma(h2, ps, -1.0 / 5040);
      ps = fma(h2, ps, 1.0 / 120);
      ps = fma(h2, ps, -1.0 / 6);
      ps = fma(h2, ps, 1.0);
      sh = h * ps;
      double pc = 1.0 / 20922789888000.0;
      pc = fma(h2, pc, -1.0 / 87178291200.0);
      pc = fma(h2, pc, 1.0 / 479001600);
      pc = fma(h2, pc, -1.0 / 3628800);
      pc = fma(h2, pc, 1.0 / 40320);
      pc = fma(h2, pc, -1.0 / 720);
      pc = fma(h2, pc, 1.0 / 24);
      pc = fma(h2, pc, -0.5);
      ch = fma(h2, pc, 1.0);
    } else {
      sincos(h, &sh, &ch);
    }
    dq[0] = ax * sh;
    dq[1] = ay * sh;
    dq[2] = az * sh;
    dq[3] = ch;
    const double il = rsq64(fma(dq[0], dq[0], fma(dq[1], dq[1], fma(dq[2], dq[2], dq[3] * dq[3]))));
    for (int a = 0; a < 4; ++a) dq[a] *= il;
  }
  const double* q = st->q;
  double nq[4] = {fma(dq[3], q[0], fma(dq[0], q[3], fma(dq[1], q[2], -dq[2] * q[1]))),
                  fma(dq[3], q[1], fma(-dq[0], q[2], fma(dq[1], q[3], dq[2] * q[0]))),
                  fma(dq[3], q[2], fma(dq[0], q[1], fma(-dq[1], q[0], dq[2] * q[3]))),
                  fma(dq[3], q[3], fma(-dq[0], q[0], fma(-dq[1], q[1], -dq[2] * q[2])))};
  const double l2 = fma(nq[0], nq[0], fma(nq[1], nq[1], fma(nq[2], nq[2], nq[3] * nq[3])));
  double Rd[3][3];
  q_to_mat64(dq, Rd);
  double nt[3];
  for (int a = 0; a < 3; ++a)
    nt[a] = fma(Rd[a][0], st->t[0], fma(Rd[a][1], st->t[1], fma(Rd[a][2], st->t[2], x[3 + a])));
  // Movement bound of the next pass: p'_new - p'_old = (dR - I)(R s + t) + dt,
  // and |dR - I| = 2 sin(angle/2) <= angle, so |dp'| <= angle (|s| + |t|) + |dt|
  // (rounded up; the f32 pose's own error is covered by the search's slack).
  {
    const double tn = sqrt(st->t[0] * st->t[0] + st->t[1] * st->t[1] + st->t[2] * st->t[2]);
    const double dn = sqrt(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
    st->move_rot = float(angle * (1.0 + 1e-6) + 1e-30);
    st->move_const = float((angle * tn + dn) * (1.0 + 1e-6) + 1e-30);
  }
  const double il = rsq64(l2);
  for (int a = 0; a < 4; ++a) st->q[a] = nq[a] * il;
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

// ---- nearest target point within dmax (exact; ties to the lower index) ----
//
// Phase 1 runs per lane in the pass's first stage: the box and empty-space
// rejects, the previous pass's correspondence as a seed (a real candidate, so it
// tightens the radius without changing the answer), and the 3x3x3 block cut to
// the seeded sphere. A query the block's clearance certifies is done. The others
// (in early passes about a third, clustered in space) go to a queue that the
// second stage drains with dynamic work distribution: groups of kNNGroup lanes
// each take one queued query and split the rows of each ring of its ball
// between them, agreeing on the radius after every ring.

struct NNQuery {
  uint64_t best_key;  // (d2 bits << 32 | original index); starts at (dmax2, ~0)
  uint32_t best_pos;  // grid position of the best, kNone if none
  float thr;          // search radius^2 = min(dmax2, best d2)
  float second;       // phase 1: smallest d2 of the other points scanned
  float budget;       // phase 1: movement before the match can change (0: none)
  int hx, hy, hz;     // query cell
  bool hard;          // phase 2 needed
  bool skip_block;    // the 3x3x3 block was scanned in phase 1
  uint32_t rows, pts; // work counters (search-statistics mode)
};

// (d2, original index) packed as one u64 key: d2 >= 0, so its bits order like
// the float, and the index breaks ties; the initial key (dmax2, ~0) admits
// exactly d2 <= dmax2 (NaN distances compare above it). With kSecond the
// smallest d2 of the other points seen is kept too (the same point seen again,
// as the seed is, does not count).
template <bool kSecond = false>
__device__ __forceinline__ void nn_consider(NNQuery& Q, float px, float py, float pz, uint32_t s,
                                            const float4& q) {
  const float d2 = dist2(px, py, pz, q.x, q.y, q.z);
  const uint64_t key = (uint64_t(__float_as_uint(d2)) << 32) | __float_as_uint(q.w);
  if constexpr (kSecond) {
    const float old_best = Q.best_pos != kNone ? Q.thr : INFINITY;
    Q.second = key < Q.best_key ? fminf(Q.second, old_best)
                                : (key == Q.best_key ? Q.second : fminf(Q.second, d2));
  }
  if (key < Q.best_key) {
    Q.best_key = key;
    Q.best_pos = s;
    Q.thr = d2;
  }
}

#ifndef SR_BLOCK_W
#define SR_BLOCK_W 8
#endif
constexpr int kBlockW = SR_BLOCK_W;  // point loads in flight per lane in the phase-1 block

// Widening of the phase-1 block scan, in cells (see nn_phase1), used once the
// pose moves every point of the source less than kGapMove cells per pass:
// earlier, budgets would be spent at once and the widening and the
// second-nearest tracking would only cost.
constexpr float kGapMargin = 0.05f;
constexpr float kGapMove = 0.01f;

// Movement budget of a query with no point within dmax, when every point is
// certified at least D away: while the query moves less than D - dmax (less a
// rounding slack), it still has none.
__device__ __forceinline__ float none_budget(float D, float dmax, float px, float py, float pz) {
  const float slack = 1e-6f * (fabsf(px) + fabsf(py) + fabsf(pz) + D + 1.0f);
  return fmaxf(D - dmax - slack, 0.0f);
}

template <bool kStats, bool kGap>
__device__ __forceinline__ void nn_phase1(const GridDesc& g, const uint32_t* __restrict__ cs,
                                          const uint8_t* __restrict__ empty,
                                          const float4* __restrict__ tp, float px, float py,
                                          float pz, float dmax, float dmax2, uint32_t prev,
                                          float margin, NNQuery& Q, SearchStats* stats) {
  Q.best_key = (uint64_t(__float_as_uint(dmax2)) << 32) | 0xFFFFFFFFull;
  Q.best_pos = kNone;
  Q.thr = dmax2;
  Q.second = INFINITY;
  Q.budget = 0.0f;
  Q.hard = false;
  Q.skip_block = false;
  Q.rows = Q.pts = 0;
  // Quick reject: query farther than dmax from the grid's box.
  {
    float bx = fmaxf(fmaxf(g.ox - px, px - (g.ox + g.h * g.dx)), 0.0f);
    float by = fmaxf(fmaxf(g.oy - py, py - (g.oy + g.h * g.dy)), 0.0f);
    float bz = fmaxf(fmaxf(g.oz - pz, pz - (g.oz + g.h * g.dz)), 0.0f);
    float lb = bx * bx + by * by + bz * bz;
    if (lb > dmax2 + 2.0f * dmax * g.margin + g.margin * g.margin) {
      if constexpr (kStats) stats_add(stats, 0, 0, 0, kRejected);
      // every point lies within the box widened by the filing margin
      if constexpr (kGap) Q.budget = none_budget(sqrtf(lb) - g.margin, dmax, px, py, pz);
      return;
    }
  }
  Q.hx = cell_coord_free(px, g.ox, g.inv_h);
  Q.hy = cell_coord_free(py, g.oy, g.inv_h);
  Q.hz = cell_coord_free(pz, g.oz, g.inv_h);
  const int hx = Q.hx, hy = Q.hy, hz = Q.hz;
  const bool inside = hx >= 0 && hx < g.dx && hy >= 0 && hy < g.dy && hz >= 0 && hz < g.dz;
  int e = 0;
  if (empty && inside) {
    // Rings closer than the empty radius hold no point, so every point lies
    // outside the box of rings 0..e-1: its clearance bounds the distance.
    e = ldg(empty + ((uint32_t(hz) * uint32_t(g.dy) + uint32_t(hy)) * uint32_t(g.dx) + uint32_t(hx)));
    const float ce = e > 0 ? ring_clearance(g, px, py, pz, hx, hy, hz, e - 1) : 0.0f;
    if (ce > dmax) {
      if constexpr (kStats) stats_add(stats, 0, 0, 0, kRejected);
      if constexpr (kGap) Q.budget = none_budget(ce, dmax, px, py, pz);
      return;
    }
  }
  if (prev != kNone) nn_consider<kGap>(Q, px, py, pz, prev, ldg(tp + prev));
  if (inside && e <= 1) {
    // Rings 0 and 1 as nine contiguous point ranges (streamed as one sequence,
    // kBlockW loads in flight), cut to the sphere of the seeded radius widened
    // by `margin`, so that a certified match also comes with a gap to every
    // other point; most queries of a converging registration end here.
    const float rclip = kGap ? __builtin_amdgcn_sqrtf(Q.thr) + margin : 0.0f;
    scan_block27_within<kBlockW, true>(g, cs, tp, px, py, pz, hx, hy, hz,
                                       kGap ? rclip * rclip : Q.thr,
                                       [&](uint32_t s, const float4& q) {
      if constexpr (kStats) ++Q.pts;
      nn_consider<kGap>(Q, px, py, pz, s, q);
    });
    Q.skip_block = true;
    const float cl = ring_clearance(g, px, py, pz, hx, hy, hz, 1);
    if (cl > 0.0f && Q.thr < cl * cl) {
      if constexpr (kStats) stats_add(stats, 1, 0, Q.pts, kBlockDone);
      if (kGap) {
        // Every other point lies at least `other` away: scanned ones at
        // sqrt(second), unscanned cells of the block beyond rclip, cells
        // outside it beyond cl. While the point moves less than half the gap
        // (less a rounding slack), its nearest point stays the same; with no
        // point within dmax, every point is an other one.
        const float other = fminf(fminf(sqrtf(Q.second), rclip), cl);
        if (Q.best_pos != kNone) {
          const float best = sqrtf(Q.thr);
          const float slack = 1e-6f * (fabsf(px) + fabsf(py) + fabsf(pz) + best + 1.0f);
          Q.budget = fmaxf(0.5f * (other - best) - slack, 0.0f);
        } else {
          Q.budget = none_budget(other, dmax, px, py, pz);
        }
      }
      return;
    }
  }
  Q.hard = true;
}

#ifndef SR_NN_GROUP
#define SR_NN_GROUP 8
#endif
constexpr int kNNGroup = SR_NN_GROUP;  // lanes per query in phase 2 (a power of two <= 64)
static_assert(kNNGroup >= 1 && kNNGroup <= 64 && (kNNGroup & (kNNGroup - 1)) == 0, "group size");
#ifndef SR_ROW_BATCH
#define SR_ROW_BATCH 1
#endif
constexpr int kRowBatch = SR_ROW_BATCH;  // rows per bound-loading batch in phase 2
#ifndef SR_STREAM_W
#define SR_STREAM_W 4
#endif
constexpr int kStreamW = SR_STREAM_W;  // point loads in flight per lane in phase 2
#ifndef SR_RING_CONCAT
#define SR_RING_CONCAT 4
#endif
// Rows per lane batched in the ring search of one-registration launches (their
// grid descriptor sits in scalar registers; with per-lane descriptors, as in
// batches, the batch's registers would spill).
constexpr int kRingConcat = SR_RING_CONCAT;


// Row t of Chebyshev ring r (r >= 1) of (y,z) rows around the query's row:
// the perimeter of the (2r+1)^2 square, each row once.
__device__ __forceinline__ void ring_row(int r, int t, int& dy, int& dz) {
  if (r == 0) {
    dy = dz = 0;
    return;
  }
  const int e = 2 * r;  // rows per side; t < 4e
  const int side = int(t >= e) + int(t >= 2 * e) + int(t >= 3 * e), o = t - side * e;
  if (side == 0) {
    dz = -r;
    dy = -r + o;
  } else if (side == 1) {
    dy = r;
    dz = -r + o;
  } else if (side == 2) {
    dz = r;
    dy = r - o;
  } else {
    dy = -r;
    dz = r - o;
  }
}

// Ring search of one query by the kNNGroup lanes of a group, which must all
// call it together. W holds the best so far (phase 1) and its radius; rings are
// centred on the query's row clamped to one row outside the grid, so a query
// beside the grid still reaches every grid row within max(dy, dz) rings (the
// gap and clearance tests use the true position). On return every lane of the
// group holds the group's best key and position.
//
// kGap (converged passes, as in phase 1): the scan radius is widened by
// `margin` beyond the best distance and the smallest distance of the other
// points is kept in W.second, so that on return every point other than the
// best is known to lie at least min(sqrt(W.second), best + margin) away (see
// hard_query for the movement budget this gives). The group's lanes scan
// disjoint rows, so the other lanes' bests count as other points too.
template <bool kStats, int kG = kNNGroup, bool kGap = false, int kRB = 0>
__device__ __forceinline__ void group_search(const GridDesc& g, const uint32_t* __restrict__ cs,
                                             const float4* __restrict__ tp, float qx, float qy,
                                             float qz, bool skip, float margin, NNQuery& W) {
  static_assert(kG >= 1 && kG <= kWave && (kG & (kG - 1)) == 0, "group size");
  const int sub = int(threadIdx.x & (kWave - 1)) % kG;
  const int hx = cell_coord_free(qx, g.ox, g.inv_h);
  const int hy = min(max(cell_coord_free(qy, g.oy, g.inv_h), -1), g.dy);
  const int hz = min(max(cell_coord_free(qz, g.oz, g.inv_h), -1), g.dz);
  const int rmax = max(max(hy, g.dy - 1 - hy), max(hz, g.dz - 1 - hz));
  // squared scan radius: the best distance, widened by margin under kGap
  auto scan2 = [&](float thr) {
    if constexpr (kGap) {
      const float r = __builtin_amdgcn_sqrtf(thr) + margin;
      return r * r;
    } else {
      return thr;
    }
  };
  for (int r = 0; r <= rmax; ++r) {
    if (r > 0) {  // W.thr is the same on the group's lanes here
      const float c = clearance_yz(g, qy, qz, hy, hz, r - 1);
      if (c > 0.0f && c * c > scan2(W.thr)) break;
    }
    const int nrows = r == 0 ? 1 : 8 * r;
    auto visit = [&](uint32_t s, const float4& q) {
      if constexpr (kStats) ++W.pts;
      nn_consider<kGap>(W, qx, qy, qz, s, q);
    };
    if (kRB > 0 && r >= 2) {
      // kRB rows per lane at a time: their range bounds in one round of
      // loads, then their points as one sequence (stream_concat), so a lane's
      // rows of a ring cost about one bound round trip and one point round
      // trip per kStreamW points instead of both per row. (Rings 0 and 1,
      // whose rows skip the 3x3x3 block, take the per-row path.)
      constexpr int kRB1 = kRB > 0 ? kRB : 1;
      for (int t0 = sub; t0 < nrows; t0 += kG * kRB1) {
        uint32_t lo[kRB1], hi[kRB1];
#pragma unroll
        for (int m = 0; m < kRB1; ++m) {
          uint32_t a = 0, b = 0;
          const int t = t0 + m * kG;
          if (t < nrows) {
            int dy, dz;
            ring_row(r, t, dy, dz);
            const int cz = hz + dz, cy = hy + dy;
            if (cz >= 0 && cz < g.dz && cy >= 0 && cy < g.dy) {
              const float gz = axis_gap(qz, g.oz, g.h, g.margin, cz);
              const float gy = axis_gap(qy, g.oy, g.h, g.margin, cy);
              const float gyz = gz * gz + gy * gy;
              const float R2 = scan2(W.thr);
              if (gyz <= R2) {
                const float rx = __builtin_amdgcn_sqrtf(R2 - gyz) + g.margin;
                const int x0 = max(cell_coord_free(qx - rx, g.ox, g.inv_h) - 1, 0);
                const int x1 = min(cell_coord_free(qx + rx, g.ox, g.inv_h) + 1, g.dx - 1);
                if (x0 <= x1) {
                  const uint32_t cbase = (uint32_t(cz) * uint32_t(g.dy) + uint32_t(cy)) * uint32_t(g.dx);
                  if constexpr (kStats) ++W.rows;
                  a = cbase + uint32_t(x0);
                  b = cbase + uint32_t(x1) + 1;
                }
              }
            }
          }
          lo[m] = ldg(cs + a);
          hi[m] = ldg(cs + b);
        }
        stream_concat<kStreamW>(lo, hi, tp, visit);
      }
    } else {
      // This lane's rows of the ring, kRowBatch at a time: the bounds of a
      // batch's ranges are loaded together and then its points streamed, so a
      // batch costs one bound round trip instead of one per row. A row whose
      // (y,z) gap exceeds the radius contributes an empty range.
      for (int t0 = sub; t0 < nrows; t0 += kG * kRowBatch) {
        uint32_t ia[2 * kRowBatch], ib[2 * kRowBatch];
#pragma unroll
        for (int m = 0; m < kRowBatch; ++m) {
          ia[2 * m] = ib[2 * m] = ia[2 * m + 1] = ib[2 * m + 1] = 0;
          const int t = t0 + m * kG;
          if (t >= nrows) continue;
          int dy, dz;
          ring_row(r, t, dy, dz);
          const int cz = hz + dz, cy = hy + dy;
          if (cz < 0 || cz >= g.dz || cy < 0 || cy >= g.dy) continue;
          const float gz = axis_gap(qz, g.oz, g.h, g.margin, cz);
          const float gy = axis_gap(qy, g.oy, g.h, g.margin, cy);
          const float gyz = gz * gz + gy * gy;
          const float R2 = scan2(W.thr);
          if (gyz > R2) continue;
          const float rx = __builtin_amdgcn_sqrtf(R2 - gyz) + g.margin;
          const int x0 = max(cell_coord_free(qx - rx, g.ox, g.inv_h) - 1, 0);
          const int x1 = min(cell_coord_free(qx + rx, g.ox, g.inv_h) + 1, g.dx - 1);
          if (x0 > x1) continue;
          const uint32_t cbase = (uint32_t(cz) * uint32_t(g.dy) + uint32_t(cy)) * uint32_t(g.dx);
          if constexpr (kStats) ++W.rows;
          if (skip && r <= 1) {
            // left and right of the block's three cells
            if (x0 <= hx - 2) {
              ia[2 * m] = cbase + uint32_t(x0);
              ib[2 * m] = cbase + uint32_t(min(x1, hx - 2)) + 1;
            }
            if (x1 >= hx + 2) {
              ia[2 * m + 1] = cbase + uint32_t(max(x0, hx + 2));
              ib[2 * m + 1] = cbase + uint32_t(x1) + 1;
            }
          } else {
            ia[2 * m] = cbase + uint32_t(x0);
            ib[2 * m] = cbase + uint32_t(x1) + 1;
          }
        }
        uint32_t lo[2 * kRowBatch], hi[2 * kRowBatch];
#pragma unroll
        for (int m = 0; m < 2 * kRowBatch; ++m) {
          lo[m] = ldg(cs + ia[m]);
          hi[m] = ldg(cs + ib[m]);
        }
        stream_ranges<kStreamW>(lo, hi, tp, visit);
      }
    }
    // the group agrees on the smallest radius before the next ring
#pragma unroll
    for (int off = 1; off < kG; off <<= 1) W.thr = fminf(W.thr, __shfl_xor(W.thr, off));
  }
  const uint64_t own_key = W.best_key;
  const bool own_point = W.best_pos != kNone;
  // group minimum of (key, position)
#pragma unroll
  for (int off = 1; off < kG; off <<= 1) {
    const uint32_t ohi = __shfl_xor(uint32_t(W.best_key >> 32), off);
    const uint32_t olo = __shfl_xor(uint32_t(W.best_key), off);
    const uint32_t opos = __shfl_xor(W.best_pos, off);
    const uint64_t ok = (uint64_t(ohi) << 32) | olo;
    if (ok < W.best_key) {
      W.best_key = ok;
      W.best_pos = opos;
    }
    if constexpr (kStats) {
      W.rows += __shfl_xor(W.rows, off);
      W.pts += __shfl_xor(W.pts, off);
    }
  }
  if constexpr (kGap) {
    // a lane's own best that lost the group minimum is another point
    if (own_point && own_key != W.best_key)
      W.second = fminf(W.second, __uint_as_float(uint32_t(own_key >> 32)));
#pragma unroll
    for (int off = 1; off < kG; off <<= 1) W.second = fminf(W.second, __shfl_xor(W.second, off));
  }
}

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

// ---- the pass in four stages ----------------------------------------------
//
// k_gicp_settle: per source point, the match the last search certified for
//   the pose's movement so far is kept; the other points are marked for the
//   search. A block whose points all keep their matches adds its terms, writes
//   its partial and flags itself done.
// k_gicp_search: per marked point, p' = R p + t and phase 1. Certified points
//   record their match; uncertified ones are queued (wave-aggregated append).
// k_gicp_hard: drains the queue, kWave/kNNGroup queries per wave, so the
//   clustered hard queries spread over the whole chip instead of lengthening
//   the waves that hold them.
// k_gicp_terms: the blocks not flagged done add their terms from the recorded
//   matches and write their partials.
// A block's terms are added in the same order whichever stage does it, so the
// sums do not depend on how the points were split between the stages.

// Logical block of this workgroup, XCD-aware: the hardware deals consecutive
// workgroup ids round robin over the 8 XCDs; give each XCD a contiguous range
// of logical blocks so neighbouring source tiles (which touch the same target
// cells) share an L2.
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t G = gridDim.x, hw = blockIdx.x;
  const uint32_t xcd = hw & 7u, q8 = G >> 3, rem = G & 7u;
  return xcd * q8 + min(xcd, rem) + (hw >> 3);
}

// The search's result is a function of the transformed point's f32
// coordinates alone (the exact nearest point, ties to the lower index), so a
// point back at the very coordinates of the search that gave its match keeps
// that match whatever its budget: in converged passes the pose moves points
// far less than an f32 ulp, and this keeps the points whose budget is nil (a
// second-nearest point closer to the nearest than the budget's rounding slack)
// from being searched on every pass. Only a point without a budget reads it,
// and only with match reuse on (off, every point is searched on every pass).
__device__ __forceinline__ bool same_spot(const float4* seen, uint32_t i, float px, float py,
                                          float pz) {
  const float4 s = ldg(seen + i);
  return s.x == px && s.y == py && s.z == pz;  // NaN (never searched) compares unequal
}

// The f32 pose of a pass: high parts R, T and low parts Rl, Tl.
struct PassPose {
  float R[9], T[3], Rl[9], Tl[3];
};

__device__ __forceinline__ void load_pose(const InstState* st, PassPose& P) {
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    P.R[a] = st->rf[a];
    P.Rl[a] = st->rlo[a];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    P.T[a] = st->tf[a];
    P.Tl[a] = st->tlo[a];
  }
}

// p' = R p + t (Isometry3::transform_point, transform.rs:126-128) with the FP64
// pose split into f32 high and low parts: the low part removes the systematic
// error of an f32-rounded pose, which would otherwise floor the Gauss-Newton
// update at ~1e-7 (the reference's own f32 pose has that floor).
__device__ __forceinline__ void transform_point(const PassPose& P, const float4& p, float& px,
                                                float& py, float& pz) {
  const float lx = fmaf(P.Rl[0], p.x, fmaf(P.Rl[1], p.y, fmaf(P.Rl[2], p.z, P.Tl[0])));
  const float ly = fmaf(P.Rl[3], p.x, fmaf(P.Rl[4], p.y, fmaf(P.Rl[5], p.z, P.Tl[1])));
  const float lz = fmaf(P.Rl[6], p.x, fmaf(P.Rl[7], p.y, fmaf(P.Rl[8], p.z, P.Tl[2])));
  px = fmaf(P.R[0], p.x, fmaf(P.R[1], p.y, fmaf(P.R[2], p.z, P.T[0] + lx)));
  py = fmaf(P.R[3], p.x, fmaf(P.R[4], p.y, fmaf(P.R[5], p.z, P.T[1] + ly)));
  pz = fmaf(P.R[6], p.x, fmaf(P.R[7], p.y, fmaf(P.R[8], p.z, P.T[2] + lz)));
}

// Adds one correspondence's H (upper triangle, row-major over the 6x6), g and
// squared error to the per-thread sums.
__device__ __forceinline__ void add_terms(float (&acc)[kTerms], const PassPose& P, float beta,
                                          float eps, float px, float py, float pz, const float4& q,
                                          const float4& nt, const float4& ns, float d2) {
  const float* R = P.R;
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
}

// Block reduction (fixed order, compensated). Rows padded by 8 floats and
// chunks read interleaved (element e*8 + ch), so the 64 lanes of a wave spread
// over the LDS banks. kRows terms go through LDS at a time (all 28, or half of
// them in two rounds, for kernels whose occupancy the LDS would limit); the
// arithmetic is the same either way.
constexpr int kRedRow = kBlock + 8;
template <int kRows>
struct RedLdsT {
  float red[kRows][kRedRow];
  float part_hi[kTerms][8], part_lo[kTerms][8];
  unsigned long long wcount[kBlock / kWave];
};
using RedLds = RedLdsT<kTerms>;
using RedLdsHalf = RedLdsT<kTerms / 2>;
static_assert(kTerms % 2 == 0, "terms split in halves");

// Writes the block partial: 28 (hi, lo) pairs + u64 count. The finalize kernel
// of this pass reads it after the kernel boundary, so no device-scope fence (an
// L2 write-back + invalidate on every XCD) is needed here.
template <int kRows>
__device__ __forceinline__ void write_partial(RedLdsT<kRows>& L, const float (&acc)[kTerms],
                                              uint32_t count, float* out) {
  unsigned long long c64 = count;
  for (int off = kWave / 2; off > 0; off >>= 1) c64 += __shfl_xor(c64, off);
  if ((threadIdx.x & (kWave - 1)) == 0) L.wcount[threadIdx.x / kWave] = c64;
#pragma unroll
  for (int r0 = 0; r0 < kTerms; r0 += kRows) {
    if (r0 > 0) __syncthreads();  // the previous round's rows have been read
#pragma unroll
    for (int t = 0; t < kRows; ++t) L.red[t][threadIdx.x] = acc[r0 + t];
    __syncthreads();
    if (threadIdx.x < kRows * 8) {
      const int t = threadIdx.x >> 3, ch = threadIdx.x & 7;
      Kahan k;
      k.init();
      for (int e = 0; e < kBlock / 8; ++e) k.add(L.red[t][e * 8 + ch]);
      L.part_hi[r0 + t][ch] = k.s;
      L.part_lo[r0 + t][ch] = -k.c;
    }
  }
  __syncthreads();
  if (threadIdx.x < kTerms) {
    const int t = threadIdx.x;
    float hi = 0.0f, lo = 0.0f;
    for (int ch = 0; ch < 8; ++ch) {
      float s, e;
      two_sum(hi, L.part_hi[t][ch], s, e);
      hi = s;
      lo += e + L.part_lo[t][ch];
    }
    float s, e;
    two_sum(hi, lo, s, e);
    out[2 * t] = s;
    out[2 * t + 1] = e;
  }
  if (threadIdx.x == kTerms) {
    unsigned long long tot = 0;
    for (int w = 0; w < kBlock / kWave; ++w) tot += L.wcount[w];
    reinterpret_cast<unsigned long long*>(out)[30] = tot;  // floats 60..61
  }
}

constexpr uint32_t kHardBlockFlag = 0x80000000u;  // HardQuery.inst: block scanned in phase 1

// The queue stage is bound by the latency of its chains of dependent loads, so
// it is compiled for as many resident waves as the SIMD takes.
#ifndef SR_HARD_WAVES
#define SR_HARD_WAVES 8
#endif
// Consecutive queue slices per XCD chunk in the queue stage (see k_gicp_hard).
constexpr uint32_t kHardChunk = 64;

// The one-kernel first stage at 3 waves per SIMD (its registers then do not
// spill; cfg1 4.18 -> 4.15 ms against 4).
#ifndef SR_SEARCH_WAVES
#define SR_SEARCH_WAVES 3
#endif
// The search stage over marked points (batches) gains from a fifth wave per
// SIMD although registers then spill: its waves mostly wait on the dependent
// loads of a few marked lanes (measured on 256 pairs: 31 passes 13.7 -> 13.3 ms;
// the one-kernel form is slower at 5).
#ifndef SR_MARKED_WAVES
#define SR_MARKED_WAVES 5
#endif

// Queues this lane's uncertified point (hard) for the ring search: one atomic
// per wave (the leader reserves the wave's slots).
__device__ __forceinline__ void enqueue_hard(const PassScratch& ps, bool hard, const NNQuery& Q,
                                             float px, float py, float pz, uint32_t inst_id,
                                             uint32_t i, float* budget) {
  const int lane = int(threadIdx.x & (kWave - 1));
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long hm = __ballot(hard);
  if (!hm) return;
  const int leader = __ffsll((long long)hm) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(&ps.ctl[0], uint32_t(__popcll(hm)));
  base = __shfl(base, leader);
  if (!hard) return;
  HardQuery h;
  h.px = px;
  h.py = py;
  h.pz = pz;
  h.inst = inst_id | (Q.skip_block ? kHardBlockFlag : 0u);
  h.i = i;
  h.pos = Q.best_pos;
  h.key = Q.best_key;
  ps.queue[base + uint32_t(__popcll(hm & below))] = h;
  budget[i] = 0.0f;
}

// enqueue_hard for kernels whose threads all call it together: one atomic
// per workgroup instead of one per wave (the queue length is a single counter
// that every workgroup of the launch bumps).
__device__ __forceinline__ void enqueue_hard_block(const PassScratch& ps, bool hard,
                                                   const NNQuery& Q, float px, float py, float pz,
                                                   uint32_t inst_id, uint32_t i, float* budget) {
  __shared__ uint32_t wcount[kBlock / kWave];
  __shared__ uint32_t bbase;
  const int lane = int(threadIdx.x & (kWave - 1)), wave = int(threadIdx.x / kWave);
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long hm = __ballot(hard);
  __syncthreads();  // a previous call's readers of wcount / bbase are done
  if (lane == 0) wcount[wave] = uint32_t(__popcll(hm));
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) tot += wcount[w];
    bbase = tot ? atomicAdd(&ps.ctl[0], tot) : 0u;
  }
  __syncthreads();
  if (!hard) return;
  uint32_t base = bbase;
  for (int w = 0; w < wave; ++w) base += wcount[w];
  HardQuery h;
  h.px = px;
  h.py = py;
  h.pz = pz;
  h.inst = inst_id | (Q.skip_block ? kHardBlockFlag : 0u);
  h.i = i;
  h.pos = Q.best_pos;
  h.key = Q.best_key;
  ps.queue[base + uint32_t(__popcll(hm & below))] = h;
  budget[i] = 0.0f;
}

// First stage. One-point-per-thread blocks: each point whose match the last
// search certified for the pose's movement so far keeps it (budget spent, or
// the match dropped when beyond d_max); the others are marked in the block's
// mask for the search stage. A block whose points all keep their matches adds
// its terms right away from two rounds of loads (its own points, then the
// matched target points) and is flagged done; the other blocks get their
// partials from k_gicp_terms. Blocks of several points per thread are left to
// the search stage whole. Compiled without the search, so many workgroups are
// resident to hide the two rounds (the block reduction goes through LDS in two
// halves for the same reason).
template <bool kStats>
__global__ void __launch_bounds__(kBlock) k_gicp_settle(const IterInstance* __restrict__ insts,
                                                        const uint32_t* __restrict__ block_inst,
                                                        PassScratch ps,
                                                        SearchStats* __restrict__ stats) {
  const uint32_t b = xcd_block();
  const IterInstance& I = insts[block_inst[b]];
  const InstState* st = I.state;
  if (st->phase == 2 || I.ppt != 1) {  // converged (every stage skips it) or searched whole
    if (threadIdx.x == 0) ps.done[b] = st->phase == 2 ? 1u : 0u;
    return;
  }
  const uint32_t lb = b - I.block_begin;
  const float move_rot = st->move_rot, move_const = st->move_const;
  const InstConfig cfg = I.cfg;
  PassPose P;
  load_pose(st, P);
  const uint32_t i = lb * kBlock + threadIdx.x;
  const bool valid = i < I.n_src;
  const float4 p = ldg(I.src_rec + kRec * (valid ? i : 0u));
  const float4 ns = ldg(I.src_rec + kRec * (valid ? i : 0u) + 1);
  const uint32_t prev = valid ? ldg(I.match + (i)) : kNone;
  const float bud = valid ? ldg(I.budget + (i)) : 0.0f;
  const float move =
      move_rot * (sqrtf(p.x * p.x + p.y * p.y + p.z * p.z) * 1.000001f) + move_const;
  float px = 0.0f, py = 0.0f, pz = 0.0f, d2 = 0.0f;
  transform_point(P, p, px, py, pz);
  // a match, or none (kNone), certified for the movement or searched right here
  const bool cert = valid && bud > move;
  const bool keep =
      cert || (valid && (cfg.flags & kCfgMatchReuse) && same_spot(I.seen, i, px, py, pz));
  const bool search = valid && !keep;
  const unsigned long long sm = __ballot(search);
  if ((threadIdx.x & (kWave - 1)) == 0) ps.pend[size_t(b) * (kBlock / kWave) + threadIdx.x / kWave] = sm;
  float4 q = p, nt = p;
  bool counted = false;
  if (keep) {
    if constexpr (kStats) stats_add(stats, 0, 0, 0, kBlockDone);
    if (prev == kNone) {
      if (cert) I.budget[i] = bud - move;
    } else {
      q = ldg(I.tgt_rec + kRec * prev);
      nt = ldg(I.tgt_rec + kRec * prev + 1);
      d2 = dist2(px, py, pz, q.x, q.y, q.z);
      if (d2 <= cfg.dmax2) {
        if (cert) I.budget[i] = bud - move;
        counted = true;
      } else {  // the nearest point is beyond d_max: none here, search next time
        I.match[i] = kNone;
        I.budget[i] = 0.0f;
        I.seen[i] = make_float4(px, py, pz, 0.0f);
      }
    }
  }
  const bool settled = __syncthreads_and(search ? 0 : 1);
  if (threadIdx.x == 0) ps.done[b] = settled ? 1u : 0u;
  if (!settled) return;
  const float beta = cfg.one_minus_eps, eps = 1.0f - cfg.one_minus_eps;
  float acc[kTerms];
#pragma unroll
  for (int t = 0; t < kTerms; ++t) acc[t] = 0.0f;
  uint32_t count = 0;
  if (counted) {
    add_terms(acc, P, beta, eps, px, py, pz, q, nt, ns, d2);
    ++count;
  }
  __shared__ RedLdsHalf L;
  write_partial(L, acc, count, I.partials + size_t(lb) * 64);
}

// One block of several points per thread: every point's match (and movement
// budget) is kept or searched anew and recorded, and the uncertified ones are
// queued; the block's terms come from k_gicp_terms.
template <bool kStats>
__device__ __forceinline__ void search_block(const IterInstance* __restrict__ insts,
                                             const uint32_t* __restrict__ block_inst,
                                             const PassScratch& ps, SearchStats* __restrict__ stats,
                                             uint32_t b) {
  const uint32_t inst_id = block_inst[b];
  const IterInstance& I = insts[inst_id];
  const InstState* st = I.state;
  if (st->phase == 2) return;
  const uint32_t lb = b - I.block_begin;
  const GridDesc g = I.grid;
  const InstConfig cfg = I.cfg;
  PassPose P;
  load_pose(st, P);
  const int lane = int(threadIdx.x & (kWave - 1));
  const unsigned long long below = (1ull << lane) - 1ull;
  const float move_rot = st->move_rot, move_const = st->move_const;
  // budgets are worth computing once no source point moves kGapMove cells
  // (and only with match reuse on: without a budget every point is searched)
  const bool gap = (cfg.flags & kCfgMatchReuse) && move_rot * I.src_radius + move_const < kGapMove * g.h;
  const uint32_t ppt = I.ppt;
  for (uint32_t k = 0; k < ppt; ++k) {
    const uint32_t i = (lb * ppt + k) * kBlock + threadIdx.x;
    const bool valid = i < I.n_src;
    const float4 p = ldg(I.src_rec + kRec * (valid ? i : 0u));
    float px, py, pz;
    transform_point(P, p, px, py, pz);
    const uint32_t prev = valid ? ldg(I.match + (i)) : kNone;
    // The last search left this point a movement budget (half its gap to the
    // second-nearest point); while the pose moves it less, the match stands.
    const float bud = valid ? ldg(I.budget + (i)) : 0.0f;
    const float move =
        move_rot * (sqrtf(p.x * p.x + p.y * p.y + p.z * p.z) * 1.000001f) + move_const;
    // a match, or none (kNone), certified for the movement or searched right here
    const bool cert = bud > move;
    const bool keep =
        cert || (valid && (cfg.flags & kCfgMatchReuse) && same_spot(I.seen, i, px, py, pz));
    NNQuery Q;
    Q.hard = false;
    if (!keep) {
      if (gap)
        nn_phase1<kStats, true>(g, I.cell_start, I.empty_radius, I.tgt, px, py, pz, cfg.dmax,
                                cfg.dmax2, prev, kGapMargin * g.h, Q, valid ? stats : nullptr);
      else
        nn_phase1<kStats, false>(g, I.cell_start, I.empty_radius, I.tgt, px, py, pz, cfg.dmax,
                                 cfg.dmax2, prev, 0.0f, Q, valid ? stats : nullptr);
    }
    const bool hard = valid && Q.hard;
    enqueue_hard_block(ps, hard, Q, px, py, pz, inst_id, i, I.budget);
    if (!valid || hard) continue;
    if (keep) {
      if constexpr (kStats) stats_add(stats, 0, 0, 0, kBlockDone);
      if (prev == kNone) {
        if (cert) I.budget[i] = bud - move;
        continue;
      }
      const float4 q = ldg(I.tgt_rec + kRec * prev);
      if (dist2(px, py, pz, q.x, q.y, q.z) <= cfg.dmax2) {
        if (cert) I.budget[i] = bud - move;
      } else {  // the nearest point is beyond d_max: none here, search next time
        I.match[i] = kNone;
        I.budget[i] = 0.0f;
        I.seen[i] = make_float4(px, py, pz, 0.0f);
      }
      continue;
    }
    I.budget[i] = Q.budget;
    I.match[i] = Q.best_pos;
    I.seen[i] = make_float4(px, py, pz, 0.0f);
  }
}

// Second stage, the blocks k_gicp_settle did not finish.
template <bool kStats>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(SR_MARKED_WAVES))) k_gicp_search(const IterInstance* __restrict__ insts,
                                                        const uint32_t* __restrict__ block_inst,
                                                        PassScratch ps,
                                                        SearchStats* __restrict__ stats) {
  const uint32_t b = xcd_block();
  if (ps.done[b]) return;
  const uint32_t inst_id = block_inst[b];
  const IterInstance& I = insts[inst_id];
  if (I.ppt != 1) {  // the whole block
    search_block<kStats>(insts, block_inst, ps, stats, b);
    return;
  }
  // The points k_gicp_settle marked, a wave per 64 threads of the block (waves
  // with none leave at once); their terms come from k_gicp_terms.
  const int lane = int(threadIdx.x & (kWave - 1));
  const unsigned long long sm = ps.pend[size_t(b) * (kBlock / kWave) + threadIdx.x / kWave];
  if (sm == 0) return;
  const bool mine = (sm >> lane) & 1ull;
  const InstState* st = I.state;
  const GridDesc g = I.grid;
  const InstConfig cfg = I.cfg;
  PassPose P;
  load_pose(st, P);
  // budgets are worth computing once no source point moves kGapMove cells
  // (and only with match reuse on: without a budget every point is searched)
  const bool gap = (cfg.flags & kCfgMatchReuse) &&
                   st->move_rot * I.src_radius + st->move_const < kGapMove * g.h;
  const uint32_t i = (b - I.block_begin) * kBlock + threadIdx.x;
  float px = 0.0f, py = 0.0f, pz = 0.0f;
  NNQuery Q;
  Q.hard = false;
  if (mine) {
    transform_point(P, ldg(I.src_rec + kRec * i), px, py, pz);
    const uint32_t prev = ldg(I.match + (i));
    if (gap)
      nn_phase1<kStats, true>(g, I.cell_start, I.empty_radius, I.tgt, px, py, pz, cfg.dmax,
                              cfg.dmax2, prev, kGapMargin * g.h, Q, stats);
    else
      nn_phase1<kStats, false>(g, I.cell_start, I.empty_radius, I.tgt, px, py, pz, cfg.dmax,
                               cfg.dmax2, prev, 0.0f, Q, stats);
  }
  const bool hard = mine && Q.hard;
  enqueue_hard(ps, hard, Q, px, py, pz, inst_id, i, I.budget);
  if (mine && !hard) {
    I.budget[i] = Q.budget;
    I.match[i] = Q.best_pos;
    I.seen[i] = make_float4(px, py, pz, 0.0f);
  }
}

// The first two stages in one kernel, for launches too small to fill the chip
// (there, an extra launch per pass costs more than the settled blocks' higher
// occupancy saves): settled one-point-per-thread blocks add their terms in one
// sweep; other blocks keep or search each point, and a block none of whose
// points was queued adds its terms from the matches just recorded. Matches,
// budgets and sums are those of the two-kernel form.
template <bool kStats>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(SR_SEARCH_WAVES))) k_gicp_search_fused(const IterInstance* __restrict__ insts,
                                                        const uint32_t* __restrict__ block_inst,
                                                        PassScratch ps,
                                                        SearchStats* __restrict__ stats) {
  const uint32_t b = xcd_block();
  const uint32_t inst_id = block_inst[b];
  const IterInstance& I = insts[inst_id];
  const InstState* st = I.state;
  if (st->phase == 2) return;
  const uint32_t lb = b - I.block_begin;
  const GridDesc g = I.grid;
  const InstConfig cfg = I.cfg;
  PassPose P;
  load_pose(st, P);
  const int lane = int(threadIdx.x & (kWave - 1));
  const unsigned long long below = (1ull << lane) - 1ull;
  const float move_rot = st->move_rot, move_const = st->move_const;
  // budgets are worth computing once no source point moves kGapMove cells
  // (and only with match reuse on: without a budget every point is searched)
  const bool gap = (cfg.flags & kCfgMatchReuse) && move_rot * I.src_radius + move_const < kGapMove * g.h;
  const uint32_t ppt = I.ppt;
  const float beta = cfg.one_minus_eps, eps = 1.0f - cfg.one_minus_eps;
  __shared__ RedLds L;

  // Settled blocks: when every point of a one-point-per-thread block keeps its
  // match, the search is skipped and the terms are added in the same sweep, so
  // the block costs two rounds of loads (its own points, then the matched
  // target points) instead of four. The terms are those loop 2 would add.
  if (ppt == 1) {
    const uint32_t i = lb * kBlock + threadIdx.x;
    const bool valid = i < I.n_src;
    const float4 p = ldg(I.src_rec + kRec * (valid ? i : 0u));
    const float4 ns = ldg(I.src_rec + kRec * (valid ? i : 0u) + 1);
    const uint32_t prev = valid ? ldg(I.match + (i)) : kNone;
    const float bud = valid ? ldg(I.budget + (i)) : 0.0f;
    const float move =
        move_rot * (sqrtf(p.x * p.x + p.y * p.y + p.z * p.z) * 1.000001f) + move_const;
    float px, py, pz;
    transform_point(P, p, px, py, pz);
    // a match, or none (kNone), certified for the movement or searched right here
    const bool cert = bud > move;
    const bool keep =
        cert || (valid && (cfg.flags & kCfgMatchReuse) && same_spot(I.seen, i, px, py, pz));
    if (__syncthreads_and(!valid || keep)) {
      if (threadIdx.x == 0) ps.done[b] = 1u;
      float acc[kTerms];
#pragma unroll
      for (int t = 0; t < kTerms; ++t) acc[t] = 0.0f;
      uint32_t count = 0;
      if (valid && prev == kNone) {
        if constexpr (kStats) stats_add(stats, 0, 0, 0, kBlockDone);
        if (cert) I.budget[i] = bud - move;
      } else if (valid) {
        if constexpr (kStats) stats_add(stats, 0, 0, 0, kBlockDone);
        const float4 q = ldg(I.tgt_rec + kRec * prev);
        const float4 nt = ldg(I.tgt_rec + kRec * prev + 1);
        const float d2 = dist2(px, py, pz, q.x, q.y, q.z);
        if (d2 <= cfg.dmax2) {
          if (cert) I.budget[i] = bud - move;
          add_terms(acc, P, beta, eps, px, py, pz, q, nt, ns, d2);
          ++count;
        } else {  // the nearest point is beyond d_max: none here
          I.match[i] = kNone;
          I.budget[i] = 0.0f;
          I.seen[i] = make_float4(px, py, pz, 0.0f);
        }
      }
      write_partial(L, acc, count, I.partials + size_t(lb) * 64);
      return;
    }
  }

  // Loop 1, the search: every point's match (and movement budget) is recorded,
  // and the uncertified ones are queued.
  bool queued = false;
  for (uint32_t k = 0; k < ppt; ++k) {
    const uint32_t i = (lb * ppt + k) * kBlock + threadIdx.x;
    const bool valid = i < I.n_src;
    const float4 p = ldg(I.src_rec + kRec * (valid ? i : 0u));
    float px, py, pz;
    transform_point(P, p, px, py, pz);
    const uint32_t prev = valid ? ldg(I.match + (i)) : kNone;
    // The last search left this point a movement budget (half its gap to the
    // second-nearest point); while the pose moves it less, the match stands.
    const float bud = valid ? ldg(I.budget + (i)) : 0.0f;
    const float move =
        move_rot * (sqrtf(p.x * p.x + p.y * p.y + p.z * p.z) * 1.000001f) + move_const;
    // a match, or none (kNone), certified for the movement or searched right here
    const bool cert = bud > move;
    const bool keep =
        cert || (valid && (cfg.flags & kCfgMatchReuse) && same_spot(I.seen, i, px, py, pz));
    NNQuery Q;
    Q.hard = false;
    if (!keep) {
      if (gap)
        nn_phase1<kStats, true>(g, I.cell_start, I.empty_radius, I.tgt, px, py, pz, cfg.dmax,
                                cfg.dmax2, prev, kGapMargin * g.h, Q, valid ? stats : nullptr);
      else
        nn_phase1<kStats, false>(g, I.cell_start, I.empty_radius, I.tgt, px, py, pz, cfg.dmax,
                                 cfg.dmax2, prev, 0.0f, Q, valid ? stats : nullptr);
    }
    const bool hard = valid && Q.hard;
    queued = queued || __ballot(hard) != 0ull;
    enqueue_hard_block(ps, hard, Q, px, py, pz, inst_id, i, I.budget);
    if (!valid || hard) continue;
    if (keep) {
      if constexpr (kStats) stats_add(stats, 0, 0, 0, kBlockDone);
      if (prev == kNone) {
        if (cert) I.budget[i] = bud - move;
        continue;
      }
      const float4 q = ldg(I.tgt_rec + kRec * prev);
      if (dist2(px, py, pz, q.x, q.y, q.z) <= cfg.dmax2) {
        if (cert) I.budget[i] = bud - move;
      } else {  // the nearest point is beyond d_max: none here, search next time
        I.match[i] = kNone;
        I.budget[i] = 0.0f;
        I.seen[i] = make_float4(px, py, pz, 0.0f);
      }
      continue;
    }
    I.budget[i] = Q.budget;
    I.match[i] = Q.best_pos;
    I.seen[i] = make_float4(px, py, pz, 0.0f);
  }
  // A block with queued points gets its partial from k_gicp_terms.
  const int any = __syncthreads_or(queued ? 1 : 0);
  if (threadIdx.x == 0) ps.done[b] = any ? 0u : 1u;
  if (any) return;
  // Loop 2, the terms, from the matches this thread just wrote (a separate
  // loop, so the search's registers and the sums' are not live together).
  float acc[kTerms];
#pragma unroll
  for (int t = 0; t < kTerms; ++t) acc[t] = 0.0f;
  uint32_t count = 0;
  for (uint32_t k = 0; k < ppt; ++k) {
    const uint32_t i = (lb * ppt + k) * kBlock + threadIdx.x;
    if (i >= I.n_src) continue;
    const uint32_t j = ldg(I.match + (i));
    if (j == kNone) continue;
    float px, py, pz;
    transform_point(P, ldg(I.src_rec + kRec * i), px, py, pz);
    const float4 q = ldg(I.tgt_rec + kRec * j);
    const float d2 = dist2(px, py, pz, q.x, q.y, q.z);
    add_terms(acc, P, beta, eps, px, py, pz, q, ldg(I.tgt_rec + kRec * j + 1),
              ldg(I.src_rec + kRec * i + 1), d2);
    ++count;
  }
  write_partial(L, acc, count, I.partials + size_t(lb) * 64);
}

// One queued point's ring search by a group of kG lanes (all of them call it).
template <bool kStats, bool kSingle, int kG>
__device__ __forceinline__ void hard_query(const IterInstance* __restrict__ insts, const HardQuery& h,
                                           bool active, SearchStats* __restrict__ stats) {
  const IterInstance& I = insts[kSingle ? 0u : (h.inst & ~kHardBlockFlag)];
  NNQuery W;
  W.best_key = h.key;
  W.best_pos = h.pos;
  W.thr = __uint_as_float(uint32_t(h.key >> 32));
  W.second = INFINITY;
  W.rows = W.pts = 0;
  if (!active) return;
  const GridDesc g = I.grid;
  const InstState* st = I.state;
  // the same test as the first stage's, on the same state: budgets are worth
  // computing once no source point moves kGapMove cells
  const bool gap = (I.cfg.flags & kCfgMatchReuse) &&
                   st->move_rot * I.src_radius + st->move_const < kGapMove * g.h;
  const float margin = kGapMargin * g.h;
  constexpr int kRB = kSingle ? kRingConcat : 0;
  if (gap)  // the 3x3x3 block is scanned again: its second-nearest counts
    group_search<kStats, kG, true, kRB>(g, I.cell_start, I.tgt, h.px, h.py, h.pz, false, margin, W);
  else
    group_search<kStats, kG, false, kRB>(g, I.cell_start, I.tgt, h.px, h.py, h.pz,
                                         (h.inst & kHardBlockFlag) != 0, 0.0f, W);
  if (int(threadIdx.x & (kWave - 1)) % kG == 0) {
    I.match[h.i] = W.best_pos;
    I.seen[h.i] = make_float4(h.px, h.py, h.pz, 0.0f);
    if (gap) {
      // Every other point lies at least `other` away: scanned ones at
      // sqrt(second), unscanned ones beyond the widened radius (rows cut to
      // it, rings stopped once their clearance exceeded it). While the point
      // moves less than half the gap (less a rounding slack), its nearest
      // point stays the same: the movement budget phase 1 gives its own
      // certified points. With none within dmax (the radius stayed dmax),
      // every point is an other one.
      const float best = sqrtf(W.thr);
      const float other = fminf(sqrtf(W.second), best + margin);
      if (W.best_pos != kNone) {
        const float slack = 1e-6f * (fabsf(h.px) + fabsf(h.py) + fabsf(h.pz) + best + 1.0f);
        I.budget[h.i] = fmaxf(0.5f * (other - best) - slack, 0.0f);
      } else {
        I.budget[h.i] = none_budget(other, best, h.px, h.py, h.pz);
      }
    }
    if constexpr (kStats)
      stats_add(stats, 2, W.rows, W.pts, W.best_pos == kNone ? kRingNone : kRingFound);
  }
}

// kSingle: the launch holds one registration, so every queued point belongs to
// insts[0] and its grid descriptor and pointers are wave-uniform (scalar
// registers) instead of per-lane copies (cfg1: 64 -> 50 VGPRs, no spill).
// Static split, kWave/kG queries per wave and round (a slice): the grid holds
// a wave per slice of a fully queued launch up to a cap, so every wave takes
// few slices (no shared cursor to contend on).
// kG: lanes per queued point (a launch-tuning choice: any group size gives the
// same matches, since every point of a ring is visited by exactly one lane and
// the group takes the minimum packed key).
template <bool kStats, bool kSingle, int kG>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(SR_HARD_WAVES))) k_gicp_hard(const IterInstance* __restrict__ insts,
                                                      PassScratch ps,
                                                      SearchStats* __restrict__ stats) {
  const uint32_t qlen = ps.ctl[0];
  const int lane = int(threadIdx.x & (kWave - 1));
  constexpr uint32_t G = kWave / kG;
  constexpr uint32_t kWavesPerBlock = kBlock / kWave;
  const uint32_t waves = gridDim.x * kWavesPerBlock;
  // Slice order. The queue is in source order and hard queries cluster in
  // space, so giving each XCD one contiguous share of it (as the other stages
  // do for locality) left some XCDs with most of the expensive slices. Here XCD
  // x (workgroup id mod 8) takes chunks x, x + 8, x + 16, ... of kHardChunk
  // consecutive slices, in dispatch order: every XCD samples the whole queue,
  // a chunk's queries still share target cells in its L2, and slices beyond
  // the resident waves go to waves dispatched as earlier ones finish. (A
  // bijection on the wave ids: hard_stage_blocks makes the grid a multiple of
  // 2 * kHardChunk workgroups.)
  const uint32_t xcd = blockIdx.x & 7u;
  const uint32_t j = (blockIdx.x >> 3) * kWavesPerBlock + threadIdx.x / kWave;
  const uint32_t wid = ((j / kHardChunk) * 8u + xcd) * kHardChunk + (j % kHardChunk);
  for (uint32_t base = wid * G; base < qlen; base += waves * G) {
    const uint32_t e = base + uint32_t(lane / kG);
    const bool active = e < qlen;  // the same on a group's lanes
    const HardQuery h = ldg(ps.queue + (active ? e : base));
    hard_query<kStats, kSingle, kG>(insts, h, active, stats);
  }
}

// A workgroup per run of consecutive logical blocks (runs kept on one XCD), so
// a pass whose blocks were all finished by k_gicp_search costs a small launch.
__global__ void __launch_bounds__(kBlock) k_gicp_terms(const IterInstance* __restrict__ insts,
                                                       const uint32_t* __restrict__ block_inst,
                                                       uint32_t total_blocks, PassScratch ps) {
  const uint32_t w = xcd_block(), G = gridDim.x;
  const uint32_t per = total_blocks / G, rem = total_blocks % G;
  const uint32_t b0 = w * per + min(w, rem), b1 = b0 + per + (w < rem ? 1u : 0u);
  __shared__ RedLds L;
  for (uint32_t b = b0; b < b1; ++b) {
    if (ps.done[b]) continue;
    const uint32_t inst_id = block_inst[b];
    const IterInstance& I = insts[inst_id];
    const InstState* st = I.state;
    if (st->phase == 2) continue;
    const uint32_t lb = b - I.block_begin;
    const InstConfig cfg = I.cfg;
    const float beta = cfg.one_minus_eps, eps = 1.0f - cfg.one_minus_eps;
    PassPose P;
    load_pose(st, P);
    float acc[kTerms];
#pragma unroll
    for (int t = 0; t < kTerms; ++t) acc[t] = 0.0f;
    uint32_t count = 0;
    const uint32_t ppt = I.ppt;
    for (uint32_t k = 0; k < ppt; ++k) {
      const uint32_t i = (lb * ppt + k) * kBlock + threadIdx.x;
      if (i >= I.n_src) continue;
      const uint32_t j = ldg(I.match + (i));
      if (j == kNone) continue;
      float px, py, pz;
      transform_point(P, ldg(I.src_rec + kRec * i), px, py, pz);
      const float4 q = ldg(I.tgt_rec + kRec * j);
      // the squared distance the search compared (same operations, same operands)
      const float d2 = dist2(px, py, pz, q.x, q.y, q.z);
      add_terms(acc, P, beta, eps, px, py, pz, q, ldg(I.tgt_rec + kRec * j + 1),
                ldg(I.src_rec + kRec * i + 1), d2);
      ++count;
    }
    write_partial(L, acc, count, I.partials + size_t(lb) * 64);
    __syncthreads();  // the LDS is reused by the next block
  }
}

// One workgroup per registration: FP64 sum of its block partials, then the
// Gauss-Newton bookkeeping, solve and pose update on one thread. The partials
// are summed in a fixed order (below), the same on every run whatever the
// launch mix.
constexpr int kFinBlock = 1024;
constexpr int kFinGroups = kFinBlock / 32;

__global__ void __launch_bounds__(kFinBlock) k_gicp_finalize(const IterInstance* __restrict__ insts,
                                                             uint32_t count, uint32_t* ctl) {
  // the pass's queue is drained: empty it for the next pass
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl[0] = 0u;
  const uint32_t inst_id = blockIdx.x;
  if (inst_id >= count) return;
  const IterInstance& I = insts[inst_id];
  InstState* st = I.state;
  const InstConfig cfg = I.cfg;
  // The state is staged in LDS: the serial bookkeeping below reads and writes
  // it field by field, which from global memory would chain many round trips.
  // Its loads and the partials' go out together; a finished registration (its
  // partials stale but readable) leaves after the first barrier.
  static_assert(sizeof(InstState) % 8 == 0 && sizeof(InstState) / 8 <= kFinBlock, "state staging");
  constexpr int kStateWords = int(sizeof(InstState) / 8);
  __shared__ InstState sst;
  if (threadIdx.x < kStateWords)
    reinterpret_cast<unsigned long long*>(&sst)[threadIdx.x] =
        ldg(reinterpret_cast<const unsigned long long*>(st) + threadIdx.x);
  // A block partial is 32 u64 slots: slot t < kTerms holds term t as an f32
  // (hi, lo) pair, slot 30 the correspondence count. Thread (group, slot)
  // sums one slot over partials group, group + kFinGroups, ... in FP64, so a
  // wave reads whole partials (coalesced) and no cross-lane reduction is needed;
  // the kFinGroups sums of a slot are then added in group order.
  __shared__ double gsum[kFinGroups][32];
  __shared__ unsigned long long gcnt[kFinGroups];
  __shared__ double dtot[kTerms];
  const uint32_t nb = I.nblocks;
  const int slot = threadIdx.x & 31, grp = threadIdx.x >> 5;
  const unsigned long long* pp = reinterpret_cast<const unsigned long long*>(I.partials);
  double acc = 0.0;
  unsigned long long cnt = 0;
  // kFinUnroll loads in flight per thread, then their adds in partial order
  constexpr int kFinUnroll = 8;  // (24, one round for 768 partials, measured slower)
  for (uint32_t b0 = grp; b0 < nb; b0 += kFinGroups * kFinUnroll) {
    unsigned long long w[kFinUnroll];
#pragma unroll
    for (int u = 0; u < kFinUnroll; ++u) {
      const uint32_t bb = b0 + uint32_t(u) * kFinGroups;
      w[u] = bb < nb ? ldg(pp + (size_t(bb) * 32 + slot)) : 0ull;
    }
#pragma unroll
    for (int u = 0; u < kFinUnroll; ++u) {
      if (slot < kTerms)
        acc += double(__uint_as_float(uint32_t(w[u]))) + double(__uint_as_float(uint32_t(w[u] >> 32)));
      else if (slot == 30)
        cnt += w[u];
    }
  }
  if (slot < kTerms) gsum[grp][slot] = acc;
  if (slot == 30) gcnt[grp] = cnt;
  __syncthreads();
  if (sst.phase == 2) return;  // the same on every thread
  if (threadIdx.x < kTerms) {
    double sacc = 0.0;
    for (int w = 0; w < kFinGroups; ++w) sacc += gsum[w][threadIdx.x];
    dtot[threadIdx.x] = sacc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long c = 0;
    for (int w = 0; w < kFinGroups; ++w) c += gcnt[w];
    finalize(&sst, cfg, dtot, dtot + 21, dtot[27], c);
  }
  __syncthreads();
  if (threadIdx.x < kStateWords)
    reinterpret_cast<unsigned long long*>(st)[threadIdx.x] =
        reinterpret_cast<const unsigned long long*>(&sst)[threadIdx.x];
}

uint32_t hard_stage_blocks(uint64_t points, uint32_t cap) {
  // a wave per kWave/kNNGroup queries of a fully queued launch, capped so that
  // an empty queue costs a small launch (the default cap, twice what the chip
  // holds at once, lets the waves beyond the resident ones pick up slices as
  // slots free; queues beyond the cap take several rounds per wave)
  const uint64_t per_block = uint64_t(kBlock / kWave) * (kWave / kNNGroup);
  const uint64_t want = (points + per_block - 1) / per_block;
  uint64_t b = std::min<uint64_t>(std::max<uint64_t>(want, 1), std::max<uint32_t>(cap, 1));
  constexpr uint64_t kRound = 2 * kHardChunk;  // the queue stage's slice order needs it
  b = (b + kRound - 1) / kRound * kRound;
  return uint32_t(b);
}

// Workgroups of the terms stage, each taking a run of blocks: several rounds
// of what 256 CUs hold at once.
#ifndef SR_TERMS_GRID
#define SR_TERMS_GRID 4096
#endif
constexpr uint32_t kTermsGrid = SR_TERMS_GRID;

template <bool kStats, bool kSingle>
static void launch_hard(int group, uint32_t hard_blocks, hipStream_t s, const IterInstance* insts,
                        const PassScratch& ps, SearchStats* st) {
  switch (group) {
    case 4:
      hipLaunchKernelGGL((k_gicp_hard<kStats, kSingle, 4>), dim3(hard_blocks), dim3(kBlock), 0, s, insts, ps, st);
      break;
    case 16:
      hipLaunchKernelGGL((k_gicp_hard<kStats, kSingle, 16>), dim3(hard_blocks), dim3(kBlock), 0, s, insts, ps, st);
      break;
    default:
      hipLaunchKernelGGL((k_gicp_hard<kStats, kSingle, kNNGroup>), dim3(hard_blocks), dim3(kBlock), 0, s, insts, ps, st);
  }
}

void gicp_pass(const IterInstance* insts, const uint32_t* block_inst, uint32_t total_blocks,
               uint32_t count, const PassScratch& ps, uint32_t hard_blocks, hipStream_t s,
               bool finalize, const PassTuning& tune) {
  if (total_blocks == 0) return;
  {
    ProfScope prof("gicp_pass", s);
    SearchStats* st = prof_stats("gicp_pass");
    // Launches of at least tune.split_blocks blocks (several per CU) run the
    // first two stages as two kernels; smaller ones as one.
    const bool split = total_blocks >= tune.split_blocks;
    if (st) {
      if (split) {
        hipLaunchKernelGGL(k_gicp_settle<true>, dim3(total_blocks), dim3(kBlock), 0, s, insts,
                           block_inst, ps, st);
        hipLaunchKernelGGL(k_gicp_search<true>, dim3(total_blocks), dim3(kBlock), 0, s, insts,
                           block_inst, ps, st);
      } else {
        hipLaunchKernelGGL(k_gicp_search_fused<true>, dim3(total_blocks), dim3(kBlock), 0, s,
                           insts, block_inst, ps, st);
      }
      if (count == 1) launch_hard<true, true>(tune.group, hard_blocks, s, insts, ps, st);
      else launch_hard<true, false>(tune.group, hard_blocks, s, insts, ps, st);
    } else {
      if (split) {
        hipLaunchKernelGGL(k_gicp_settle<false>, dim3(total_blocks), dim3(kBlock), 0, s, insts,
                           block_inst, ps, st);
        hipLaunchKernelGGL(k_gicp_search<false>, dim3(total_blocks), dim3(kBlock), 0, s, insts,
                           block_inst, ps, st);
      } else {
        hipLaunchKernelGGL(k_gicp_search_fused<false>, dim3(total_blocks), dim3(kBlock), 0, s,
                           insts, block_inst, ps, st);
      }
      if (count == 1) launch_hard<false, true>(tune.group, hard_blocks, s, insts, ps, st);
      else launch_hard<false, false>(tune.group, hard_blocks, s, insts, ps, st);
    }
    hipLaunchKernelGGL(k_gicp_terms, dim3(std::min<uint32_t>(total_blocks, kTermsGrid)), dim3(kBlock),
                       0, s, insts, block_inst, total_blocks, ps);
  }
  if (!finalize) return;
  ProfScope prof("gicp_finalize", s);
  hipLaunchKernelGGL(k_gicp_finalize, dim3(count), dim3(kFinBlock), 0, s, insts, count, ps.ctl);
}

}  // namespace sr
