// CPU oracle for the GICP scan-registration path — TEST INFRASTRUCTURE ONLY.
//
// This file restates, in C++, the algorithm of SpatialRust's CPU GICP path so the
// HIP implementation can be checked against it. It is never linked into, called
// by, or used as a fallback of the product library; only tests/, the smoke check
// and the CPU-baseline leg of bench.py load it.
//
// Two precisions share one template:
//   faithful (P = float): f32 exactly where the reference uses f32 (positions,
//     the KD-tree, squared distances, the Isometry3<f32> pose, f32 products in
//     the covariance accumulator) and f64 where it uses f64 (Jacobi, GICP normal
//     equations). Built with -ffp-contract=off so no FMA changes a rounding.
//   fp64 (P = double): the same algorithm with every quantity in double — the
//     "FP64 CPU oracle" the GPU pose is compared against.
//
// Reference files followed (paths relative to the rsasaki0109/SpatialRust checkout):
//   crates/spatialrust-registration/src/gicp.rs        GICP driver and helpers
//   crates/spatialrust-search/src/kdtree.rs            KD-tree build + k-NN
//   crates/spatialrust-search/src/brute.rs             brute-force k-NN
//   crates/spatialrust-search/src/uniform_grid.rs      grid_bounds / build_grid
//   crates/spatialrust-math/src/{covariance,eigen,linalg,quat,transform,mat,vec}.rs
//   crates/spatialrust-filtering/src/voxel.rs          centroid voxel downsample
//   crates/spatialrust-gpu/src/kernels/covariances_grid.rs  radius covariances
//   crates/spatialrust-gpu/src/kernels/normals_grid.rs      radius normals
//
// Pinning status of each function is recorded in the tests that exercise it
// (tests/test_oracle_*.py); parity-unpinned functions are listed there.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <vector>
#include <omp.h>

#include "../include/sr_gicp.h"

namespace {

// ---------------------------------------------------------------- math ----

template <typename T>
struct V3 {
  T x, y, z;
};

template <typename T>
struct Q4 {
  T x, y, z, w;
};

template <typename T>
struct M3 {
  T m[3][3];
};

// Rust's f32::signum / f64::signum: 1 for +0 and positives, -1 for -0 and negatives.
inline double signum(double v) { return std::copysign(1.0, v); }

// Vec3::normalize (crates/spatialrust-math/src/vec.rs:94-100).
template <typename T>
V3<T> v_normalize(V3<T> v) {
  T len = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  if (len == T(0)) return v;
  return {v.x / len, v.y / len, v.z / len};
}

// Quat::normalize (crates/spatialrust-math/src/quat.rs:24-32).
template <typename T>
Q4<T> q_normalize(Q4<T> q) {
  T len = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  if (len == T(0)) return {T(0), T(0), T(0), T(1)};
  return {q.x / len, q.y / len, q.z / len, q.w / len};
}

// Quat::from_axis_angle (quat.rs:44-49).
template <typename T>
Q4<T> q_from_axis_angle(V3<T> axis, T angle) {
  axis = v_normalize(axis);
  T half = angle * T(0.5);
  T s = std::sin(half);
  return q_normalize(Q4<T>{axis.x * s, axis.y * s, axis.z * s, std::cos(half)});
}

// Quat::to_mat3 (quat.rs:53-71): rows built then transposed.
template <typename T>
M3<T> q_to_mat3(Q4<T> q0) {
  Q4<T> q = q_normalize(q0);
  T xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
  T xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
  T wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
  T r[3][3] = {{T(1) - T(2) * (yy + zz), T(2) * (xy + wz), T(2) * (xz - wy)},
               {T(2) * (xy - wz), T(1) - T(2) * (xx + zz), T(2) * (yz + wx)},
               {T(2) * (xz + wy), T(2) * (yz - wx), T(1) - T(2) * (xx + yy)}};
  M3<T> out;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) out.m[a][b] = r[b][a];
  return out;
}

// Quat::mul, Hamilton product then normalize (quat.rs:76-84).
template <typename T>
Q4<T> q_mul(Q4<T> a, Q4<T> b) {
  return q_normalize(Q4<T>{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                           a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                           a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
                           a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z});
}

// Mat3::mul_vec3 (crates/spatialrust-math/src/mat.rs:46-52).
template <typename T>
V3<T> m_mul_vec(const M3<T>& m, V3<T> v) {
  return {m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z,
          m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z,
          m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z};
}

// Isometry3 (crates/spatialrust-math/src/transform.rs:498-560).
template <typename T>
struct Iso {
  Q4<T> q;
  V3<T> t;
};

template <typename T>
Iso<T> iso_compose(const Iso<T>& a, const Iso<T>& b) {  // transform.rs:104-108 (a ∘ b)
  Iso<T> out;
  out.q = q_mul(a.q, b.q);
  V3<T> rt = m_mul_vec(q_to_mat3(a.q), b.t);
  out.t = {rt.x + a.t.x, rt.y + a.t.y, rt.z + a.t.z};
  return out;
}

template <typename T>
V3<T> iso_apply(const Iso<T>& iso, V3<T> p) {  // transform.rs:126-128
  V3<T> r = m_mul_vec(q_to_mat3(iso.q), p);
  return {r.x + iso.t.x, r.y + iso.t.y, r.z + iso.t.z};
}

// symmetric_eigen3: cyclic-max Jacobi, <= 32 rotations, ascending sort
// (crates/spatialrust-math/src/eigen.rs:16-84). Eigenvectors are columns.
void symmetric_eigen3(const double in[3][3], double evals[3], double evecs[3][3]) {
  double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  std::memcpy(a, in, sizeof(a));
  for (int it = 0; it < 32; ++it) {
    int p = 0, q = 1;
    double mx = std::fabs(a[0][1]);
    if (std::fabs(a[0][2]) > mx) {
      mx = std::fabs(a[0][2]);
      p = 0;
      q = 2;
    }
    if (std::fabs(a[1][2]) > mx) {
      mx = std::fabs(a[1][2]);
      p = 1;
      q = 2;
    }
    if (mx < 1e-12) break;
    double app = a[p][p], aqq = a[q][q], apq = a[p][q];
    double phi = 0.5 * (aqq - app) / apq;
    double t = signum(phi) / (std::fabs(phi) + std::sqrt(1.0 + phi * phi));
    double c = 1.0 / std::sqrt(1.0 + t * t);
    double s = t * c;
    double app_new = c * c * app - 2.0 * s * c * apq + s * s * aqq;
    double aqq_new = s * s * app + 2.0 * s * c * apq + c * c * aqq;
    a[p][p] = app_new;
    a[q][q] = aqq_new;
    a[p][q] = 0.0;
    a[q][p] = 0.0;
    for (int r = 0; r < 3; ++r) {
      if (r != p && r != q) {
        double arp = a[r][p], arq = a[r][q];
        a[r][p] = c * arp - s * arq;
        a[p][r] = a[r][p];
        a[r][q] = s * arp + c * arq;
        a[q][r] = a[r][q];
      }
    }
    for (int r = 0; r < 3; ++r) {
      double vrp = v[r][p], vrq = v[r][q];
      v[r][p] = c * vrp - s * vrq;
      v[r][q] = s * vrp + c * vrq;
    }
  }
  double ev[3] = {a[0][0], a[1][1], a[2][2]};
  int order[3] = {0, 1, 2};
  std::stable_sort(order, order + 3, [&](int i, int j) { return ev[i] < ev[j]; });
  for (int k = 0; k < 3; ++k) {
    evals[k] = ev[order[k]];
    for (int r = 0; r < 3; ++r) evecs[r][k] = v[r][order[k]];
  }
}

// solve_linear_system: Gaussian elimination with partial pivoting, pivot
// |a| < 1e-12 => singular (crates/spatialrust-math/src/linalg.rs:12-53).
bool solve_linear_system(std::vector<std::vector<double>> a, std::vector<double> b,
                         std::vector<double>& x) {
  size_t n = b.size();
  if (n == 0 || a.size() != n) return false;
  for (size_t col = 0; col < n; ++col) {
    size_t piv = col;
    for (size_t row = col + 1; row < n; ++row)
      if (std::fabs(a[row][col]) > std::fabs(a[piv][col])) piv = row;
    if (std::fabs(a[piv][col]) < 1e-12) return false;
    if (piv != col) {
      std::swap(a[piv], a[col]);
      std::swap(b[piv], b[col]);
    }
    for (size_t row = col + 1; row < n; ++row) {
      double f = a[row][col] / a[col][col];
      for (size_t k = col; k < n; ++k) a[row][k] -= f * a[col][k];
      b[row] -= f * b[col];
    }
  }
  x.assign(n, 0.0);
  for (size_t row = n; row-- > 0;) {
    double sum = b[row];
    for (size_t col = row + 1; col < n; ++col) sum -= a[row][col] * x[col];
    x[row] = sum / a[row][row];
  }
  return true;
}

// ---------------------------------------------------------- KD-tree -----

constexpr size_t kLeafSize = 16;  // kdtree.rs:10
constexpr uint8_t kAxisLeaf = 255;
constexpr uint32_t kInvalid = 0xFFFFFFFFu;

template <typename P>
struct Neighbor {
  uint32_t index;
  P d2;
};

// KdTree (crates/spatialrust-search/src/kdtree.rs:17-54, 416-610).
template <typename P>
struct KdTree {
  struct Node {
    P split;
    uint32_t left, right, start, end;
    uint8_t axis;
  };
  std::vector<P> x, y, z;
  std::vector<uint32_t> order;
  std::vector<Node> nodes;
  uint32_t root = kInvalid;

  P coord(uint32_t i, uint8_t axis) const { return axis == 0 ? x[i] : (axis == 1 ? y[i] : z[i]); }

  KdTree(const P* px, const P* py, const P* pz, size_t n) : x(px, px + n), y(py, py + n), z(pz, pz + n) {
    order.resize(n);
    for (size_t i = 0; i < n; ++i) order[i] = uint32_t(i);
    nodes.reserve(n == 0 ? 0 : ((n + kLeafSize - 1) / kLeafSize) * 2);
    if (n > 0) root = build(0, n);
  }

  uint8_t select_axis(size_t start, size_t end) const {  // kdtree.rs:458-487
    P mn[3] = {std::numeric_limits<P>::infinity(), std::numeric_limits<P>::infinity(),
               std::numeric_limits<P>::infinity()};
    P mx[3] = {-std::numeric_limits<P>::infinity(), -std::numeric_limits<P>::infinity(),
               -std::numeric_limits<P>::infinity()};
    for (size_t k = start; k < end; ++k) {
      uint32_t i = order[k];
      mn[0] = std::fmin(mn[0], x[i]);
      mn[1] = std::fmin(mn[1], y[i]);
      mn[2] = std::fmin(mn[2], z[i]);
      mx[0] = std::fmax(mx[0], x[i]);
      mx[1] = std::fmax(mx[1], y[i]);
      mx[2] = std::fmax(mx[2], z[i]);
    }
    uint8_t best = 0;
    P best_ext = mx[0] - mn[0];
    for (uint8_t a = 1; a < 3; ++a) {
      P e = mx[a] - mn[a];
      if (e > best_ext) {
        best_ext = e;
        best = a;
      }
    }
    return best;
  }

  size_t partition(size_t start, size_t end, uint8_t axis) {  // kdtree.rs:511-534
    size_t pivot_index = (start + end) / 2;
    std::swap(order[start], order[pivot_index]);
    P pv = coord(order[start], axis);
    size_t store = start + 1;
    for (size_t i = start + 1; i < end; ++i) {
      if (coord(order[i], axis) < pv) {
        std::swap(order[i], order[store]);
        ++store;
      }
    }
    std::swap(order[start], order[store - 1]);
    return store - 1;
  }

  void select_nth(size_t start, size_t end, uint8_t axis, size_t nth) {  // kdtree.rs:489-509
    size_t left = start, right = end;
    while (left < right) {
      size_t pivot = partition(left, right, axis);
      if (nth < pivot)
        right = pivot;
      else if (nth > pivot)
        left = pivot + 1;
      else
        break;
    }
  }

  uint32_t build(size_t start, size_t end) {  // kdtree.rs:416-456
    uint32_t idx = uint32_t(nodes.size());
    nodes.push_back(Node{P(0), kInvalid, kInvalid, uint32_t(start), uint32_t(end), 0});
    size_t count = end - start;
    if (count <= kLeafSize) {
      nodes[idx].axis = kAxisLeaf;
      return idx;
    }
    uint8_t axis = select_axis(start, end);
    size_t mid = start + count / 2;
    select_nth(start, end, axis, mid);
    nodes[idx].axis = axis;
    nodes[idx].split = coord(order[mid], axis);
    uint32_t l = build(start, mid);
    uint32_t r = build(mid, end);
    nodes[idx].left = l;
    nodes[idx].right = r;
    return idx;
  }

  // KnnAccumulator (kdtree.rs:551-610).
  struct Acc {
    std::vector<Neighbor<P>>& nb;
    size_t worst_index = 0;
    P worst_d2 = P(0);
    explicit Acc(std::vector<Neighbor<P>>& v) : nb(v) {}
    P prune(size_t k) const { return nb.size() < k ? std::numeric_limits<P>::infinity() : worst_d2; }
    void insert(size_t k, Neighbor<P> c) {
      if (k == 0) return;
      if (nb.size() < k) {
        nb.push_back(c);
        if (nb.size() == 1 || c.d2 > worst_d2) {
          worst_index = nb.size() - 1;
          worst_d2 = c.d2;
        }
        return;
      }
      if (c.d2 >= worst_d2) return;
      nb[worst_index] = c;
      size_t wi = 0;
      P wd = nb[0].d2;
      for (size_t i = 1; i < nb.size(); ++i)
        if (nb[i].d2 > wd) {
          wi = i;
          wd = nb[i].d2;
        }
      worst_index = wi;
      worst_d2 = wd;
    }
  };

  void recurse(uint32_t node, P qx, P qy, P qz, size_t k, Acc& best) const {  // kdtree.rs:73-121
    if (node == kInvalid) return;
    const Node& nd = nodes[node];
    if (nd.axis == kAxisLeaf) {
      for (uint32_t o = nd.start; o < nd.end; ++o) {
        uint32_t i = order[o];
        P dx = x[i] - qx, dy = y[i] - qy, dz = z[i] - qz;  // squared_distance, kdtree.rs:544-549
        best.insert(k, Neighbor<P>{i, dx * dx + dy * dy + dz * dz});
      }
      return;
    }
    P diff = (nd.axis == 0 ? qx : (nd.axis == 1 ? qy : qz)) - nd.split;
    uint32_t near = diff <= P(0) ? nd.left : nd.right;
    uint32_t far = diff <= P(0) ? nd.right : nd.left;
    recurse(near, qx, qy, qz, k, best);
    P worst = best.prune(k);
    if (diff * diff < worst || best.nb.size() < k) recurse(far, qx, qy, qz, k, best);
  }

  // nearest_k: unsorted search then a stable sort by distance (kdtree.rs:310-346).
  void nearest_k(P qx, P qy, P qz, size_t k, std::vector<Neighbor<P>>& out) const {
    out.clear();
    if (x.empty() || k == 0) return;
    Acc acc(out);
    recurse(root, qx, qy, qz, k, acc);
    std::stable_sort(out.begin(), out.end(),
                     [](const Neighbor<P>& a, const Neighbor<P>& b) { return a.d2 < b.d2; });
  }
};

// ------------------------------------------------------- covariances ----

// CovarianceAccumulator3 (crates/spatialrust-math/src/covariance.rs:25-73): f64
// sums of the inputs and of their products taken in the input precision P.
template <typename P>
struct CovAcc {
  uint64_t count = 0;
  double sum[3] = {0, 0, 0};
  double sq[6] = {0, 0, 0, 0, 0, 0};
  void push(P x, P y, P z) {
    ++count;
    sum[0] += double(x);
    sum[1] += double(y);
    sum[2] += double(z);
    sq[0] += double(P(x * x));
    sq[1] += double(P(y * y));
    sq[2] += double(P(z * z));
    sq[3] += double(P(x * y));
    sq[4] += double(P(x * z));
    sq[5] += double(P(y * z));
  }
  bool covariance(double c[3][3]) const {
    if (count < 2) return false;
    double n = double(count);
    double mx = sum[0] / n, my = sum[1] / n, mz = sum[2] / n;
    double inv = 1.0 / (n - 1.0);
    double c00 = inv * (sq[0] - n * mx * mx);
    double c11 = inv * (sq[1] - n * my * my);
    double c22 = inv * (sq[2] - n * mz * mz);
    double c01 = inv * (sq[3] - n * mx * my);
    double c02 = inv * (sq[4] - n * mx * mz);
    double c12 = inv * (sq[5] - n * my * mz);
    double m[3][3] = {{c00, c01, c02}, {c01, c11, c12}, {c02, c12, c22}};
    std::memcpy(c, m, sizeof(m));
    return true;
  }
};

// regularize: eigenvalues replaced by (epsilon, 1, 1) (gicp.rs:277-291).
void regularize(const double cov[3][3], double eps, double out[3][3], double normal[3]) {
  double ev[3], v[3][3];
  symmetric_eigen3(cov, ev, v);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) out[r][c] = 0.0;
  const double scales[3] = {eps, 1.0, 1.0};
  for (int col = 0; col < 3; ++col) {
    double axis[3] = {v[0][col], v[1][col], v[2][col]};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) out[r][c] += scales[col] * axis[r] * axis[c];
  }
  if (normal) {
    normal[0] = v[0][0];
    normal[1] = v[1][0];
    normal[2] = v[2][0];
  }
}

// covariances(): per point, k-NN (self included) -> accumulator -> regularize
// (gicp.rs:259-274). Count < 2 gives the identity before regularization.
// threads > 1 (the informational all-core timing only) splits the points over
// OpenMP threads; each point's result is the same either way.
template <typename P>
void point_covariances(const KdTree<P>& tree, size_t k, double eps, std::vector<double>& out9,
                       double* normals, int threads = 1) {
  size_t n = tree.x.size();
  out9.assign(n * 9, 0.0);
#pragma omp parallel num_threads(threads) if (threads > 1)
  {
  std::vector<Neighbor<P>> nb;
  nb.reserve(k);
#pragma omp for schedule(dynamic, 512)
  for (size_t i = 0; i < n; ++i) {
    tree.nearest_k(tree.x[i], tree.y[i], tree.z[i], k, nb);
    CovAcc<P> acc;
    for (auto& e : nb) acc.push(tree.x[e.index], tree.y[e.index], tree.z[e.index]);
    double c[3][3];
    if (!acc.covariance(c)) {
      double id[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
      std::memcpy(c, id, sizeof(c));
    }
    double r[3][3];
    regularize(c, eps, r, normals ? normals + 3 * i : nullptr);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) out9[i * 9 + a * 3 + b] = r[a][b];
  }
  }
}

// ------------------------------------------------------------- GICP -----

typedef double Mat[3][3];

inline void mat_mul(const Mat a, const Mat b, Mat out) {  // gicp.rs:379-389
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += a[r][k] * b[k][c];
      out[r][c] = s;
    }
}

inline bool inverse3(const Mat a, Mat out) {  // gicp.rs:422-441
  double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
               a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
               a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  if (std::fabs(det) < 1e-18) return false;
  double inv = 1.0 / det;
  out[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) * inv;
  out[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * inv;
  out[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * inv;
  out[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * inv;
  out[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * inv;
  out[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * inv;
  out[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) * inv;
  out[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * inv;
  out[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * inv;
  return true;
}

struct IterTrace {
  double q[4], t[3];
  double fitness;  // fitness under the pose after this iteration (NaN if not evaluated)
  double update_norm;
  uint64_t count;
};

// euclidean_fitness (gicp.rs:330-357): mean squared inlier distance (f64 of
// the P-precision differences); f64::MAX when there are no inliers.
// threads > 1 (all-core timing only): per-thread sums added in thread order.
template <typename P>
double euclidean_fitness(const std::vector<V3<P>>& pts, const KdTree<P>& tree, P maxd2,
                         std::vector<Neighbor<P>>& scratch, int threads = 1) {
  auto one = [&](const V3<P>& p, std::vector<Neighbor<P>>& nb, double& sum, uint64_t& count) {
    tree.nearest_k(p.x, p.y, p.z, 1, nb);
    if (nb.empty()) return;
    if (nb[0].d2 > maxd2) return;
    uint32_t j = nb[0].index;
    double dx = double(P(p.x - tree.x[j]));
    double dy = double(P(p.y - tree.y[j]));
    double dz = double(P(p.z - tree.z[j]));
    sum += dx * dx + dy * dy + dz * dz;
    ++count;
  };
  double sum = 0.0;
  uint64_t count = 0;
  if (threads <= 1) {
    for (auto& p : pts) one(p, scratch, sum, count);
  } else {
    std::vector<double> ts(size_t(threads), 0.0);
    std::vector<uint64_t> tc(size_t(threads), 0);
#pragma omp parallel num_threads(threads)
    {
      std::vector<Neighbor<P>> nb;
      double ls = 0.0;
      uint64_t lc = 0;
#pragma omp for schedule(static)
      for (size_t i = 0; i < pts.size(); ++i) one(pts[i], nb, ls, lc);
      ts[size_t(omp_get_thread_num())] = ls;
      tc[size_t(omp_get_thread_num())] = lc;
    }
    for (int t = 0; t < threads; ++t) {
      sum += ts[size_t(t)];
      count += tc[size_t(t)];
    }
  }
  if (count == 0) return std::numeric_limits<double>::max();
  return sum / double(count);
}

struct AlignOut {
  int32_t status;  // 0 ok, 3 invalid argument
  int32_t error_kind;  // 1 empty, 2 k<3, 3 too few correspondences, 4 singular
  uint64_t error_count;
  double q[4], t[3];
  double fitness;
  uint64_t iterations;
  uint8_t converged;
};

// GicpRegistration::align_with_diagnostics (gicp.rs:103-224).
template <typename P>
AlignOut gicp_align(const float* sx, const float* sy, const float* sz, size_t ns, const float* tx,
                    const float* ty, const float* tz, size_t nt, const sr_gicp_config& cfg,
                    const double* src_cov9_in, const double* tgt_cov9_in, IterTrace* trace,
                    size_t trace_cap, int threads = 1) {
  AlignOut out{};
  if (ns == 0 || nt == 0) {
    out.status = SR_ERR_INVALID_ARGUMENT;
    out.error_kind = 1;
    return out;
  }
  if (cfg.k_neighbors < 3) {
    out.status = SR_ERR_INVALID_ARGUMENT;
    out.error_kind = 2;
    return out;
  }
  std::vector<P> Sx(sx, sx + ns), Sy(sy, sy + ns), Sz(sz, sz + ns);
  std::vector<P> Tx(tx, tx + nt), Ty(ty, ty + nt), Tz(tz, tz + nt);
  KdTree<P> target_tree(Tx.data(), Ty.data(), Tz.data(), nt);
  std::vector<double> scov, tcov;
  if (src_cov9_in) {
    scov.assign(src_cov9_in, src_cov9_in + 9 * ns);
  } else {
    KdTree<P> st(Sx.data(), Sy.data(), Sz.data(), ns);
    point_covariances(st, cfg.k_neighbors, cfg.epsilon, scov, nullptr, threads);
  }
  if (tgt_cov9_in) {
    tcov.assign(tgt_cov9_in, tgt_cov9_in + 9 * nt);
  } else {
    KdTree<P> tt(Tx.data(), Ty.data(), Tz.data(), nt);  // the reference builds a second target tree
    point_covariances(tt, cfg.k_neighbors, cfg.epsilon, tcov, nullptr, threads);
  }

  const P dmax = P(cfg.max_correspondence_distance);
  const P maxd2 = dmax * dmax;
  Iso<P> T{{P(cfg.init_q[0]), P(cfg.init_q[1]), P(cfg.init_q[2]), P(cfg.init_q[3])},
           {P(cfg.init_t[0]), P(cfg.init_t[1]), P(cfg.init_t[2])}};
  std::vector<V3<P>> tr(ns);
  auto apply = [&]() {  // apply_transform, gicp.rs:302-312
    for (size_t i = 0; i < ns; ++i) tr[i] = iso_apply(T, V3<P>{Sx[i], Sy[i], Sz[i]});
  };
  apply();
  std::vector<Neighbor<P>> nb;
  uint64_t iterations = 0;
  bool converged = false;
  for (uint32_t it = 0; it < cfg.max_iterations; ++it) {
    ++iterations;
    M3<P> rf = q_to_mat3(T.q);
    Mat rot, rot_t;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) rot[a][b] = double(rf.m[a][b]);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) rot_t[a][b] = rot[b][a];
    double H[6][6] = {};
    double g[6] = {};
    uint64_t count = 0;
    // One correspondence's terms (gicp.rs:144-178) added to (H, g, count).
    auto body = [&](size_t i, std::vector<Neighbor<P>>& nb, double (&H)[6][6], double (&g)[6],
                    uint64_t& count) {
      const V3<P>& p = tr[i];
      target_tree.nearest_k(p.x, p.y, p.z, 1, nb);
      if (nb.empty()) return;
      if (nb[0].d2 > maxd2) return;
      uint32_t j = nb[0].index;
      Mat cs, tmp, rs, comb, m;
      std::memcpy(cs, &scov[9 * i], sizeof(cs));
      mat_mul(rot, cs, tmp);
      mat_mul(tmp, rot_t, rs);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) comb[a][b] = tcov[9 * size_t(j) + a * 3 + b] + rs[a][b];
      if (!inverse3(comb, m)) return;
      double e[3] = {double(p.x) - double(Tx[j]), double(p.y) - double(Ty[j]),
                     double(p.z) - double(Tz[j])};
      double px = double(p.x), py = double(p.y), pz = double(p.z);
      double jac[3][6] = {{0.0, pz, -py, 1.0, 0.0, 0.0},
                          {-pz, 0.0, px, 0.0, 1.0, 0.0},
                          {py, -px, 0.0, 0.0, 0.0, 1.0}};  // gicp.rs:293-300
      double mj[3][6];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 6; ++c) {
          double s = 0.0;
          for (int k = 0; k < 3; ++k) s += m[r][k] * jac[k][c];
          mj[r][c] = s;
        }
      double me[3] = {m[0][0] * e[0] + m[0][1] * e[1] + m[0][2] * e[2],
                      m[1][0] * e[0] + m[1][1] * e[1] + m[1][2] * e[2],
                      m[2][0] * e[0] + m[2][1] * e[1] + m[2][2] * e[2]};
      for (int a = 0; a < 6; ++a)
        for (int k = 0; k < 3; ++k) {
          g[a] += jac[k][a] * me[k];
          for (int b = 0; b < 6; ++b) H[a][b] += jac[k][a] * mj[k][b];
        }
      ++count;
    };
    if (threads <= 1) {
      for (size_t i = 0; i < ns; ++i) body(i, nb, H, g, count);
    } else {  // all-core timing only: per-thread sums added in thread order
      std::vector<double> th(size_t(threads) * 42, 0.0);
      std::vector<uint64_t> tcnt(size_t(threads), 0);
#pragma omp parallel num_threads(threads)
      {
        std::vector<Neighbor<P>> lnb;
        double lH[6][6] = {}, lg[6] = {};
        uint64_t lc = 0;
#pragma omp for schedule(static)
        for (size_t i = 0; i < ns; ++i) body(i, lnb, lH, lg, lc);
        double* o = th.data() + 42 * size_t(omp_get_thread_num());
        std::memcpy(o, lH, sizeof(lH));
        std::memcpy(o + 36, lg, sizeof(lg));
        tcnt[size_t(omp_get_thread_num())] = lc;
      }
      for (int t = 0; t < threads; ++t) {
        const double* o = th.data() + 42 * size_t(t);
        for (int a = 0; a < 36; ++a) H[a / 6][a % 6] += o[a];
        for (int a = 0; a < 6; ++a) g[a] += o[36 + a];
        count += tcnt[size_t(t)];
      }
    }
    if (count < cfg.min_correspondences) {
      out.status = SR_ERR_INVALID_ARGUMENT;
      out.error_kind = 3;
      out.error_count = count;
      return out;
    }
    std::vector<std::vector<double>> A(6, std::vector<double>(6));
    std::vector<double> b(6), x;
    for (int a = 0; a < 6; ++a) {
      for (int c = 0; c < 6; ++c) A[a][c] = H[a][c];
      b[a] = -g[a];
    }
    if (!solve_linear_system(A, b, x)) {
      out.status = SR_ERR_INVALID_ARGUMENT;
      out.error_kind = 4;
      return out;
    }
    // delta_from_solution (gicp.rs:314-324).
    double angle = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    Q4<P> dq{P(0), P(0), P(0), P(1)};
    if (angle > 1e-12) {
      V3<P> axis{P(x[0] / angle), P(x[1] / angle), P(x[2] / angle)};
      dq = q_from_axis_angle(axis, P(angle));
    }
    Iso<P> delta{dq, {P(x[3]), P(x[4]), P(x[5])}};
    T = iso_compose(delta, T);
    apply();
    double upd = 0.0;  // update_magnitude, gicp.rs:326-328
    for (int a = 0; a < 6; ++a) upd += x[a] * x[a];
    upd = std::sqrt(upd);
    IterTrace* tr_it = (trace && it < trace_cap) ? &trace[it] : nullptr;
    if (tr_it) {
      tr_it->q[0] = T.q.x;
      tr_it->q[1] = T.q.y;
      tr_it->q[2] = T.q.z;
      tr_it->q[3] = T.q.w;
      tr_it->t[0] = T.t.x;
      tr_it->t[1] = T.t.y;
      tr_it->t[2] = T.t.z;
      tr_it->update_norm = upd;
      tr_it->count = count;
      tr_it->fitness = std::numeric_limits<double>::quiet_NaN();
    }
    if (upd < cfg.transformation_epsilon) {
      converged = true;
      break;
    }
    double fit = euclidean_fitness(tr, target_tree, maxd2, nb, threads);
    if (tr_it) tr_it->fitness = fit;
    if (fit < cfg.fitness_epsilon) {
      converged = true;
      break;
    }
  }
  out.status = 0;
  out.q[0] = T.q.x;
  out.q[1] = T.q.y;
  out.q[2] = T.q.z;
  out.q[3] = T.q.w;
  out.t[0] = T.t.x;
  out.t[1] = T.t.y;
  out.t[2] = T.t.z;
  out.fitness = euclidean_fitness(tr, target_tree, maxd2, nb);
  out.iterations = iterations;
  out.converged = converged ? 1 : 0;
  return out;
}

// ------------------------------------------------------------ voxel -----

// VoxelGridDownsample, centroid + plain xyz schema
// (crates/spatialrust-filtering/src/voxel.rs:251-368, 550-636, 1067-1121):
// cells in first-occurrence order, f32 sequential sums, centroid = sum * (1/count).
struct VoxelCell {
  float sx, sy, sz;
  uint32_t count;
  int64_t k[3];
};

struct KeyHash3 {
  size_t operator()(const std::array<int64_t, 3>& k) const {
    uint64_t h = 0;
    for (int a = 0; a < 3; ++a) {
      h = ((h << 5) | (h >> 59)) ^ uint64_t(k[a]);
      h *= 0x517cc1b727220a95ull;
    }
    return size_t(h);
  }
};

// Rust `f32 as u32` / `as i64`: saturating, NaN -> 0.
inline uint32_t f32_as_u32(float v) {
  if (!(v == v)) return 0;
  if (v <= 0.0f) return 0;
  if (v >= 4294967296.0f) return 0xFFFFFFFFu;
  return uint32_t(v);
}
inline int64_t f32_as_i64(float v) {
  if (!(v == v)) return 0;
  if (v >= 9223372036854775808.0f) return std::numeric_limits<int64_t>::max();
  if (v < -9223372036854775808.0f) return std::numeric_limits<int64_t>::min();
  return int64_t(v);
}
}  // namespace

// ---------------------------------------------------------- exports -----

extern "C" {

int32_t or_voxel_centroid(const float* x, const float* y, const float* z, int64_t n, float leaf,
                          const float* origin, float* ox, float* oy, float* oz, int64_t* okeys,
                          uint32_t* ocounts, int64_t* n_out) {
  *n_out = 0;
  if (n == 0) return 0;  // voxel.rs:257-259: empty input returned as is
  if (leaf <= 0.0f) return SR_ERR_INVALID_ARGUMENT;  // voxel.rs:260-264
  float inv_leaf = 1.0f / leaf;
  float org[3];
  bool origin_is_min = origin == nullptr;
  bool u32_keys = false;
  if (origin) {
    org[0] = origin[0];
    org[1] = origin[1];
    org[2] = origin[2];
  } else {  // compute_bounds + fits_u32_voxel_key (voxel.rs:1067-1086)
    float mn[3] = {x[0], y[0], z[0]}, mx[3] = {x[0], y[0], z[0]};
    for (int64_t i = 1; i < n; ++i) {
      mn[0] = std::fmin(mn[0], x[i]);
      mn[1] = std::fmin(mn[1], y[i]);
      mn[2] = std::fmin(mn[2], z[i]);
      mx[0] = std::fmax(mx[0], x[i]);
      mx[1] = std::fmax(mx[1], y[i]);
      mx[2] = std::fmax(mx[2], z[i]);
    }
    const float limit = 4294967296.0f;  // u32::MAX as f32
    u32_keys = ((mx[0] - mn[0]) * inv_leaf) <= limit && ((mx[1] - mn[1]) * inv_leaf) <= limit &&
               ((mx[2] - mn[2]) * inv_leaf) <= limit;
    std::memcpy(org, mn, sizeof(org));
  }
  std::unordered_map<std::array<int64_t, 3>, uint32_t, KeyHash3> map;
  map.reserve(size_t(std::min<int64_t>(std::max<int64_t>(n / 2, 16), 1048576)));
  std::vector<VoxelCell> cells;
  for (int64_t i = 0; i < n; ++i) {
    float fx = (x[i] - org[0]) * inv_leaf;
    float fy = (y[i] - org[1]) * inv_leaf;
    float fz = (z[i] - org[2]) * inv_leaf;
    std::array<int64_t, 3> key;
    if (u32_keys) {
      key = {int64_t(f32_as_u32(fx)), int64_t(f32_as_u32(fy)), int64_t(f32_as_u32(fz))};
    } else if (origin_is_min) {
      key = {f32_as_i64(fx), f32_as_i64(fy), f32_as_i64(fz)};
    } else {
      key = {f32_as_i64(std::floor(fx)), f32_as_i64(std::floor(fy)), f32_as_i64(std::floor(fz))};
    }
    auto it = map.find(key);
    uint32_t id;
    if (it == map.end()) {
      id = uint32_t(cells.size());
      map.emplace(key, id);
      cells.push_back(VoxelCell{0.0f, 0.0f, 0.0f, 0u, {key[0], key[1], key[2]}});
    } else {
      id = it->second;
    }
    VoxelCell& c = cells[id];
    c.sx += x[i];
    c.sy += y[i];
    c.sz += z[i];
    c.count += 1;
  }
  for (size_t c = 0; c < cells.size(); ++c) {
    float inv = 1.0f / float(cells[c].count);
    ox[c] = cells[c].sx * inv;
    oy[c] = cells[c].sy * inv;
    oz[c] = cells[c].sz * inv;
    if (okeys) {
      okeys[3 * c + 0] = cells[c].k[0];
      okeys[3 * c + 1] = cells[c].k[1];
      okeys[3 * c + 2] = cells[c].k[2];
    }
    if (ocounts) ocounts[c] = cells[c].count;
  }
  *n_out = int64_t(cells.size());
  return 0;
}

// Brute-force k-NN, stable sort by f32 distance (crates/spatialrust-search/src/brute.rs:84-110).
int64_t or_brute_knn(const float* x, const float* y, const float* z, int64_t n, float qx, float qy,
                     float qz, int64_t k, int64_t* idx, float* d2) {
  if (k <= 0 || n == 0) return 0;
  std::vector<Neighbor<float>> all(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    float dx = x[i] - qx, dy = y[i] - qy, dz = z[i] - qz;
    all[size_t(i)] = Neighbor<float>{uint32_t(i), dx * dx + dy * dy + dz * dz};
  }
  std::stable_sort(all.begin(), all.end(),
                   [](const Neighbor<float>& a, const Neighbor<float>& b) { return a.d2 < b.d2; });
  int64_t m = std::min<int64_t>(k, n);
  for (int64_t i = 0; i < m; ++i) {
    idx[i] = all[size_t(i)].index;
    d2[i] = all[size_t(i)].d2;
  }
  return m;
}

// KD-tree handle (faithful f32 mode only: the reference tree is f32).
void* or_kdtree_new(const float* x, const float* y, const float* z, int64_t n) {
  return new KdTree<float>(x, y, z, size_t(n));
}
void or_kdtree_free(void* t) { delete static_cast<KdTree<float>*>(t); }
int64_t or_kdtree_nearest_k(void* t, float qx, float qy, float qz, int64_t k, int64_t* idx,
                            float* d2) {
  std::vector<Neighbor<float>> nb;
  static_cast<KdTree<float>*>(t)->nearest_k(qx, qy, qz, size_t(k), nb);
  for (size_t i = 0; i < nb.size(); ++i) {
    idx[i] = nb[i].index;
    d2[i] = nb[i].d2;
  }
  return int64_t(nb.size());
}
// Batched k-NN through the KD-tree (queries in f32), output padded with -1.
void or_kdtree_knn_batch(void* t, const float* qx, const float* qy, const float* qz, int64_t nq,
                         int64_t k, int64_t* idx, float* d2) {
  auto* tree = static_cast<KdTree<float>*>(t);
  std::vector<Neighbor<float>> nb;
  for (int64_t q = 0; q < nq; ++q) {
    tree->nearest_k(qx[q], qy[q], qz[q], size_t(k), nb);
    for (int64_t i = 0; i < k; ++i) {
      idx[q * k + i] = i < int64_t(nb.size()) ? int64_t(nb[size_t(i)].index) : -1;
      d2[q * k + i] = i < int64_t(nb.size()) ? nb[size_t(i)].d2 : -1.0f;
    }
  }
}

// Plane-regularized k-NN covariances (gicp.rs:259-291). mode 0 faithful, 1 fp64.
// out9: 9 f64 per point; normals (nullable): 3 f64 per point (column 0 of the
// sorted eigenvectors).
void or_covariances(const float* x, const float* y, const float* z, int64_t n, int64_t k,
                    double eps, int32_t mode, double* out9, double* normals) {
  std::vector<double> cov;
  if (mode == 0) {
    KdTree<float> tree(x, y, z, size_t(n));
    point_covariances(tree, size_t(k), eps, cov, normals);
  } else {
    std::vector<double> dx(x, x + n), dy(y, y + n), dz(z, z + n);
    KdTree<double> tree(dx.data(), dy.data(), dz.data(), size_t(n));
    point_covariances(tree, size_t(k), eps, cov, normals);
  }
  std::memcpy(out9, cov.data(), sizeof(double) * 9 * size_t(n));
}

void or_symmetric_eigen3(const double* m9, double* evals, double* evecs9) {
  double m[3][3], v[3][3];
  std::memcpy(m, m9, sizeof(m));
  symmetric_eigen3(m, evals, v);
  std::memcpy(evecs9, v, sizeof(v));
}

int32_t or_solve_linear_system(const double* a, const double* b, int32_t n, double* x) {
  std::vector<std::vector<double>> A(static_cast<size_t>(n), std::vector<double>(static_cast<size_t>(n)));
  std::vector<double> B(b, b + n), X;
  for (int32_t r = 0; r < n; ++r)
    for (int32_t c = 0; c < n; ++c) A[size_t(r)][size_t(c)] = a[r * n + c];
  if (!solve_linear_system(A, B, X)) return 1;
  std::memcpy(x, X.data(), sizeof(double) * size_t(n));
  return 0;
}

void or_quat_from_axis_angle(float ax, float ay, float az, float angle, float* q4) {
  Q4<float> q = q_from_axis_angle(V3<float>{ax, ay, az}, angle);
  q4[0] = q.x;
  q4[1] = q.y;
  q4[2] = q.z;
  q4[3] = q.w;
}

void or_quat_to_mat3(const float* q4, float* m9) {
  M3<float> m = q_to_mat3(Q4<float>{q4[0], q4[1], q4[2], q4[3]});
  std::memcpy(m9, m.m, sizeof(m.m));
}

// a ∘ b for isometries given as (q[4], t[3]) in f32.
void or_isometry_compose(const float* aq, const float* at, const float* bq, const float* bt,
                         float* oq, float* ot) {
  Iso<float> a{{aq[0], aq[1], aq[2], aq[3]}, {at[0], at[1], at[2]}};
  Iso<float> b{{bq[0], bq[1], bq[2], bq[3]}, {bt[0], bt[1], bt[2]}};
  Iso<float> c = iso_compose(a, b);
  oq[0] = c.q.x;
  oq[1] = c.q.y;
  oq[2] = c.q.z;
  oq[3] = c.q.w;
  ot[0] = c.t.x;
  ot[1] = c.t.y;
  ot[2] = c.t.z;
}

// Applies an f32 isometry to every point (transform_point_cloud,
// crates/spatialrust-registration/src/transform.rs:8-25).
void or_transform_points(const float* q, const float* t, const float* x, const float* y,
                         const float* z, int64_t n, float* ox, float* oy, float* oz) {
  Iso<float> iso{{q[0], q[1], q[2], q[3]}, {t[0], t[1], t[2]}};
  for (int64_t i = 0; i < n; ++i) {
    V3<float> p = iso_apply(iso, V3<float>{x[i], y[i], z[i]});
    ox[i] = p.x;
    oy[i] = p.y;
    oz[i] = p.z;
  }
}

// Result record of or_gicp_align.
typedef struct {
  int32_t status;
  int32_t error_kind;
  uint64_t error_count;
  double q[4];
  double t[3];
  double fitness;
  uint64_t iterations;
  uint8_t converged;
} or_align_result;

// GicpRegistration::align_with_diagnostics. mode 0 faithful (f32 pose/tree),
// 1 fp64. src_cov9/tgt_cov9 (nullable) replace the k-NN covariance step (used to
// feed radius covariances, the `covariance_radius` branch of gicp.rs:91-100).
// trace (nullable) receives per-iteration records (10 doubles + count each:
// q[4], t[3], fitness, update_norm, count).
int32_t or_gicp_align(const float* sx, const float* sy, const float* sz, int64_t ns,
                      const float* tx, const float* ty, const float* tz, int64_t nt,
                      const sr_gicp_config* cfg, int32_t mode, const double* src_cov9,
                      const double* tgt_cov9, or_align_result* res, double* trace,
                      int64_t trace_cap) {
  std::vector<IterTrace> tr(size_t(std::max<int64_t>(trace_cap, 0)));
  AlignOut o = mode == 0 ? gicp_align<float>(sx, sy, sz, size_t(ns), tx, ty, tz, size_t(nt), *cfg,
                                             src_cov9, tgt_cov9, tr.data(), tr.size())
                         : gicp_align<double>(sx, sy, sz, size_t(ns), tx, ty, tz, size_t(nt), *cfg,
                                              src_cov9, tgt_cov9, tr.data(), tr.size());
  res->status = o.status;
  res->error_kind = o.error_kind;
  res->error_count = o.error_count;
  std::memcpy(res->q, o.q, sizeof(o.q));
  std::memcpy(res->t, o.t, sizeof(o.t));
  res->fitness = o.fitness;
  res->iterations = o.iterations;
  res->converged = o.converged;
  if (trace) {
    for (size_t i = 0; i < tr.size() && i < o.iterations; ++i) {
      double* d = trace + 10 * i;
      std::memcpy(d, tr[i].q, 4 * sizeof(double));
      std::memcpy(d + 4, tr[i].t, 3 * sizeof(double));
      d[7] = tr[i].fitness;
      d[8] = tr[i].update_norm;
      d[9] = double(tr[i].count);
    }
  }
  return o.status;
}

// or_gicp_align in faithful precision with the per-point loops (covariances,
// correspondences, fitness) split over `threads` OpenMP threads. The
// reference is single-threaded; this exists only for an informational
// all-core CPU timing (its sums are added in another order).
int32_t or_gicp_align_mt(const float* sx, const float* sy, const float* sz, int64_t ns,
                         const float* tx, const float* ty, const float* tz, int64_t nt,
                         const sr_gicp_config* cfg, int32_t threads, or_align_result* res) {
  AlignOut o = gicp_align<float>(sx, sy, sz, size_t(ns), tx, ty, tz, size_t(nt), *cfg, nullptr,
                                 nullptr, nullptr, 0, std::max(1, int(threads)));
  res->status = o.status;
  res->error_kind = o.error_kind;
  res->error_count = o.error_count;
  std::memcpy(res->q, o.q, sizeof(o.q));
  std::memcpy(res->t, o.t, sizeof(o.t));
  res->fitness = o.fitness;
  res->iterations = o.iterations;
  res->converged = o.converged;
  return o.status;
}

// One pass of the GICP loop body (gicp.rs:137-179) in FP64 at a given pose:
// transformed points, nearest target point (KD-tree), gate, M, H and g. Writes
// H (36), g (6), the correspondence index per source point (-1 if none) and
// returns the count. Used to check H and g against finite differences.
int64_t or_gicp_normal_equations(const float* sx, const float* sy, const float* sz, int64_t ns,
                                 const float* tx, const float* ty, const float* tz, int64_t nt,
                                 const double* src_cov9, const double* tgt_cov9, const double* q4,
                                 const double* t3, double dmax, double* H36, double* g6,
                                 int64_t* corr) {
  std::vector<double> Tx(tx, tx + nt), Ty(ty, ty + nt), Tz(tz, tz + nt);
  KdTree<double> tree(Tx.data(), Ty.data(), Tz.data(), size_t(nt));
  Q4<double> q{q4[0], q4[1], q4[2], q4[3]};
  M3<double> rm = q_to_mat3(q);
  Mat rot, rot_t;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      rot[a][b] = rm.m[a][b];
      rot_t[b][a] = rm.m[a][b];
    }
  double H[6][6] = {}, g[6] = {};
  int64_t count = 0;
  std::vector<Neighbor<double>> nb;
  for (int64_t i = 0; i < ns; ++i) {
    V3<double> p = m_mul_vec(rm, V3<double>{double(sx[i]), double(sy[i]), double(sz[i])});
    p = {p.x + t3[0], p.y + t3[1], p.z + t3[2]};
    corr[i] = -1;
    tree.nearest_k(p.x, p.y, p.z, 1, nb);
    if (nb.empty() || nb[0].d2 > dmax * dmax) continue;
    uint32_t j = nb[0].index;
    Mat cs, tmp, rs, comb, m;
    std::memcpy(cs, &src_cov9[9 * i], sizeof(cs));
    mat_mul(rot, cs, tmp);
    mat_mul(tmp, rot_t, rs);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) comb[a][b] = tgt_cov9[9 * size_t(j) + a * 3 + b] + rs[a][b];
    if (!inverse3(comb, m)) continue;
    double e[3] = {p.x - Tx[j], p.y - Ty[j], p.z - Tz[j]};
    double jac[3][6] = {{0.0, p.z, -p.y, 1.0, 0.0, 0.0},
                        {-p.z, 0.0, p.x, 0.0, 1.0, 0.0},
                        {p.y, -p.x, 0.0, 0.0, 0.0, 1.0}};
    double mj[3][6];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 6; ++c) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += m[r][k] * jac[k][c];
        mj[r][c] = s;
      }
    double me[3] = {m[0][0] * e[0] + m[0][1] * e[1] + m[0][2] * e[2],
                    m[1][0] * e[0] + m[1][1] * e[1] + m[1][2] * e[2],
                    m[2][0] * e[0] + m[2][1] * e[1] + m[2][2] * e[2]};
    for (int a = 0; a < 6; ++a)
      for (int k = 0; k < 3; ++k) {
        g[a] += jac[k][a] * me[k];
        for (int b = 0; b < 6; ++b) H[a][b] += jac[k][a] * mj[k][b];
      }
    corr[i] = j;
    ++count;
  }
  std::memcpy(H36, H, sizeof(H));
  std::memcpy(g6, g, sizeof(g));
  return count;
}

// grid_bounds (crates/spatialrust-search/src/uniform_grid.rs:9-36). Returns 0 or
// SR_ERR_INVALID_ARGUMENT when the dense grid would exceed 64M cells.
int32_t or_grid_bounds(const float* x, const float* y, const float* z, int64_t n, float cell,
                       float* origin, uint32_t* dims) {
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; ++i) {
    mn[0] = std::fmin(mn[0], x[i]);
    mn[1] = std::fmin(mn[1], y[i]);
    mn[2] = std::fmin(mn[2], z[i]);
    mx[0] = std::fmax(mx[0], x[i]);
    mx[1] = std::fmax(mx[1], y[i]);
    mx[2] = std::fmax(mx[2], z[i]);
  }
  float inv = 1.0f / cell;
  for (int a = 0; a < 3; ++a) {
    int64_t span = f32_as_i64(std::floor((mx[a] - mn[a]) * inv)) + 1;
    dims[a] = uint32_t(std::max<int64_t>(span, 1));
    origin[a] = mn[a];
  }
  uint64_t cells = uint64_t(dims[0]) * dims[1] * dims[2];
  return cells > 64000000ull ? SR_ERR_INVALID_ARGUMENT : 0;
}

// Radius-grid plane covariances with the WGSL kernel's semantics, restated on the
// CPU in f32 (crates/spatialrust-gpu/src/kernels/covariances_grid.rs:81-206):
// 27-cell gather of d2 <= r2 neighbours (self included), f32 two-pass mean and
// centred sums, 16 cyclic Jacobi sweeps in f32, eigenvalues (eps,1,1) with the
// smallest mapped to eps; count < 3 gives eps*I. Output 6 f32 per point
// [c00,c11,c22,c01,c02,c12]; normals (nullable) 3 f32 per point (zero when count < 3).
// orient_normal_towards_viewpoint (crates/spatialrust-features/src/normal.rs:406-421):
// flip when n . (viewpoint - p) < 0, then Vec3::normalize (vec.rs:94-100).
static void orient_towards(float* nv, float px, float py, float pz, const float* vp) {
  float vx = vp[0] - px, vy = vp[1] - py, vz = vp[2] - pz;
  if (nv[0] * vx + nv[1] * vy + nv[2] * vz < 0.0f) {
    nv[0] = -nv[0];
    nv[1] = -nv[1];
    nv[2] = -nv[2];
  }
  float len = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
  if (len == 0.0f) return;
  nv[0] /= len;
  nv[1] /= len;
  nv[2] /= len;
}

// Shared radius-grid evaluation. out6/normals: COV_GRID_WGSL outputs
// (covariances_grid.rs:81-206); grid4: NORMALS_GRID_WGSL outputs
// (normals_grid.rs:25-198: unit normal of the smallest eigenvalue and its
// curvature, (0,0,1,0) below 3 neighbours), oriented to `vp` when given as
// GpuNormalEstimator::estimate does (features/src/normal_gpu.rs:62-71).
static int32_t radius_grid_eval(const float* x, const float* y, const float* z, int64_t n,
                                float radius, float eps, float* out6, float* normals,
                                float* grid4, const float* vp) {
  if (n == 0) return 0;
  float org[3];
  uint32_t dims[3];
  int32_t st = or_grid_bounds(x, y, z, n, radius, org, dims);
  if (st) return st;
  float inv = 1.0f / radius;
  size_t cells = size_t(dims[0]) * dims[1] * dims[2];
  auto cell_coord = [&](float v, int a) -> int64_t {
    int64_t c = f32_as_i64(std::floor((v - org[a]) * inv));
    return std::min<int64_t>(std::max<int64_t>(c, 0), int64_t(dims[a]) - 1);
  };
  // build_grid (uniform_grid.rs:44-87): counting sort, ascending index within a cell.
  std::vector<uint32_t> start(cells + 1, 0);
  std::vector<uint32_t> sorted(static_cast<size_t>(n));
  std::vector<size_t> cid(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    cid[size_t(i)] = (size_t(cell_coord(z[i], 2)) * dims[1] + size_t(cell_coord(y[i], 1))) * dims[0] +
                     size_t(cell_coord(x[i], 0));
    start[cid[size_t(i)]]++;
  }
  uint32_t acc = 0;
  for (size_t c = 0; c <= cells; ++c) {
    uint32_t v = start[c];
    start[c] = acc;
    acc += v;
  }
  std::vector<uint32_t> cursor(start.begin(), start.end() - 1);
  for (int64_t i = 0; i < n; ++i) sorted[cursor[cid[size_t(i)]]++] = uint32_t(i);
  const float r2 = radius * radius;
  for (int64_t i = 0; i < n; ++i) {
    float px = x[i], py = y[i], pz = z[i];
    int64_t c[3] = {cell_coord(px, 0), cell_coord(py, 1), cell_coord(pz, 2)};
    float mean[3] = {0, 0, 0};
    float count = 0.0f;
    for (int pass = 0; pass < 2; ++pass) {
      float c00 = 0, c11 = 0, c22 = 0, c01 = 0, c02 = 0, c12 = 0;
      for (int dz = -1; dz <= 1; ++dz) {
        int64_t nz = c[2] + dz;
        if (nz < 0 || nz >= int64_t(dims[2])) continue;
        for (int dy = -1; dy <= 1; ++dy) {
          int64_t ny = c[1] + dy;
          if (ny < 0 || ny >= int64_t(dims[1])) continue;
          for (int dx = -1; dx <= 1; ++dx) {
            int64_t nx = c[0] + dx;
            if (nx < 0 || nx >= int64_t(dims[0])) continue;
            size_t cc = (size_t(nz) * dims[1] + size_t(ny)) * dims[0] + size_t(nx);
            for (uint32_t s = start[cc]; s < start[cc + 1]; ++s) {
              uint32_t j = sorted[s];
              float ddx = x[j] - px, ddy = y[j] - py, ddz = z[j] - pz;
              if (ddx * ddx + ddy * ddy + ddz * ddz <= r2) {
                if (pass == 0) {
                  mean[0] += x[j];
                  mean[1] += y[j];
                  mean[2] += z[j];
                  count += 1.0f;
                } else {
                  float ex = x[j] - mean[0], ey = y[j] - mean[1], ez = z[j] - mean[2];
                  c00 += ex * ex;
                  c11 += ey * ey;
                  c22 += ez * ez;
                  c01 += ex * ey;
                  c02 += ex * ez;
                  c12 += ey * ez;
                }
              }
            }
          }
        }
      }
      if (pass == 0) {
        if (count < 3.0f) break;
        mean[0] /= count;
        mean[1] /= count;
        mean[2] /= count;
        continue;
      }
      float a[3][3] = {{c00, c01, c02}, {c01, c11, c12}, {c02, c12, c22}};
      float v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
      auto rotate = [&](int p, int q) {  // covariances_grid.rs:43-79
        float apq = a[p][q];
        if (std::fabs(apq) < 1e-20f) return;
        float app = a[p][p], aqq = a[q][q];
        float phi = 0.5f * (aqq - app) / apq;
        float t = phi >= 0.0f ? 1.0f / (phi + std::sqrt(1.0f + phi * phi))
                              : -1.0f / (-phi + std::sqrt(1.0f + phi * phi));
        float cs = 1.0f / std::sqrt(1.0f + t * t);
        float sn = t * cs;
        for (int r = 0; r < 3; ++r) {
          float arp = a[r][p], arq = a[r][q];
          a[r][p] = cs * arp - sn * arq;
          a[r][q] = sn * arp + cs * arq;
        }
        for (int r = 0; r < 3; ++r) {
          float apr = a[p][r], aqr = a[q][r];
          a[p][r] = cs * apr - sn * aqr;
          a[q][r] = sn * apr + cs * aqr;
        }
        for (int r = 0; r < 3; ++r) {
          float vrp = v[r][p], vrq = v[r][q];
          v[r][p] = cs * vrp - sn * vrq;
          v[r][q] = sn * vrp + cs * vrq;
        }
      };
      for (int sweep = 0; sweep < 16; ++sweep) {
        rotate(0, 1);
        rotate(0, 2);
        rotate(1, 2);
      }
      float eig[3] = {a[0][0], a[1][1], a[2][2]};
      int mi = 0;
      if (eig[1] < eig[mi]) mi = 1;
      if (eig[2] < eig[mi]) mi = 2;
      float reg[6] = {0, 0, 0, 0, 0, 0};
      for (int col = 0; col < 3; ++col) {
        float lam = col == mi ? eps : 1.0f;
        float ax = v[0][col], ay = v[1][col], az = v[2][col];
        reg[0] += lam * ax * ax;
        reg[1] += lam * ay * ay;
        reg[2] += lam * az * az;
        reg[3] += lam * ax * ay;
        reg[4] += lam * ax * az;
        reg[5] += lam * ay * az;
      }
      if (out6) std::memcpy(out6 + 6 * i, reg, sizeof(reg));
      if (normals) {
        normals[3 * i + 0] = v[0][mi];
        normals[3 * i + 1] = v[1][mi];
        normals[3 * i + 2] = v[2][mi];
      }
      if (grid4) {
        float nv[3] = {v[0][mi], v[1][mi], v[2][mi]};
        float len = std::max(std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]), 1e-20f);
        nv[0] /= len;
        nv[1] /= len;
        nv[2] /= len;
        float tr = eig[0] + eig[1] + eig[2];
        float curv = tr > 0.0f ? eig[mi] / tr : 0.0f;
        if (vp) orient_towards(nv, px, py, pz, vp);
        grid4[4 * i + 0] = nv[0];
        grid4[4 * i + 1] = nv[1];
        grid4[4 * i + 2] = nv[2];
        grid4[4 * i + 3] = curv;
      }
    }
    if (count < 3.0f) {
      float iso[6] = {eps, eps, eps, 0, 0, 0};
      if (out6) std::memcpy(out6 + 6 * i, iso, sizeof(iso));
      if (normals) normals[3 * i] = normals[3 * i + 1] = normals[3 * i + 2] = 0.0f;
      if (grid4) {
        float nv[3] = {0.0f, 0.0f, 1.0f};
        if (vp) orient_towards(nv, px, py, pz, vp);
        grid4[4 * i + 0] = nv[0];
        grid4[4 * i + 1] = nv[1];
        grid4[4 * i + 2] = nv[2];
        grid4[4 * i + 3] = 0.0f;
      }
    }
  }
  return 0;
}

int32_t or_radius_covariances(const float* x, const float* y, const float* z, int64_t n,
                              float radius, float eps, float* out6, float* normals) {
  return radius_grid_eval(x, y, z, n, radius, eps, out6, normals, nullptr, nullptr);
}

// GpuNormalEstimator::estimate with a search radius (crates/spatialrust-features/src/normal_gpu.rs:33-82)
// over NORMALS_GRID_WGSL (crates/spatialrust-gpu/src/kernels/normals_grid.rs:25-198): out4
// per point = (unit normal, curvature); vp (3 floats) may be null (no orientation).
int32_t or_grid_normals(const float* x, const float* y, const float* z, int64_t n, float radius,
                        const float* vp, float* out4) {
  return radius_grid_eval(x, y, z, n, radius, 0.0f, nullptr, nullptr, out4, vp);
}

}  // extern "C"
