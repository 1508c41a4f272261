"""Python mirror of the reference's GICP / voxel / policy surface over the C ABI.

Names, defaults, argument meaning and error messages follow
  crates/spatialrust-registration/src/gicp.rs:16-235      GicpConfig, GicpRegistration
  crates/spatialrust-registration/src/registration.rs:5-24 PointCloudRegistration, RegistrationResult
  crates/spatialrust-core/src/execution.rs:5-184           ExecutionPolicy, ExecutionReceipt
  crates/spatialrust-core/src/device.rs:4-26               DeviceKind (+ Hip)
  crates/spatialrust-filtering/src/voxel.rs:113-237        VoxelGridDownsample(Config)
  crates/spatialrust-py/src/lib.rs:2103-2121               register_gicp
Every call runs the HIP library; nothing here computes a result on the CPU.
"""
import ctypes
import enum
from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np

from . import native as _n


class SpatialError(Exception):
    """SpatialError (crates/spatialrust-core/src/error.rs:8-37)."""


class InvalidArgument(SpatialError):
    pass


class DeviceUnavailable(SpatialError):
    """No usable HIP device (an Auto policy may fall back to a CPU path)."""


class DeviceKind(enum.IntEnum):
    CPU = _n.SR_DEV_CPU
    WGPU = _n.SR_DEV_WGPU
    CUDA = _n.SR_DEV_CUDA
    HIP = _n.SR_DEV_HIP

    def is_gpu(self) -> bool:  # device.rs:22-25, extended with Hip
        return self in (DeviceKind.WGPU, DeviceKind.CUDA, DeviceKind.HIP)


@dataclass(frozen=True)
class ExecutionPolicy:
    """CpuSingle / CpuParallel / Gpu(kind) / Auto."""
    kind: int
    device: DeviceKind = DeviceKind.CPU

    @staticmethod
    def cpu_single():
        return ExecutionPolicy(_n.SR_CPU_SINGLE)

    @staticmethod
    def cpu_parallel():
        return ExecutionPolicy(_n.SR_CPU_PARALLEL)

    @staticmethod
    def gpu(device: DeviceKind = DeviceKind.HIP):
        return ExecutionPolicy(_n.SR_GPU, DeviceKind(device))

    @staticmethod
    def auto():
        return ExecutionPolicy(_n.SR_AUTO)

    def allows_fallback(self) -> bool:  # execution.rs:46-48
        return self.kind == _n.SR_AUTO

    def _c(self):
        return _n.PolicyC(int(self.kind), int(self.device))


_STAGE_NAMES = [(_n.SR_STAGE_VOXEL, "gicp-voxel"), (_n.SR_STAGE_GRID, "gicp-grid"),
                (_n.SR_STAGE_COVARIANCE, "gicp-covariance"), (_n.SR_STAGE_ITERATE, "gicp-iterate"),
                (_n.SR_STAGE_READBACK, "gpu-readback")]


@dataclass
class ExecutionReceipt:
    requested_policy: ExecutionPolicy
    resolved_policy: ExecutionPolicy
    host_to_device_bytes: int
    device_to_device_bytes: int
    device_to_host_bytes: int
    stages: Tuple[str, ...]

    @staticmethod
    def _from_c(r):
        return ExecutionReceipt(
            ExecutionPolicy(r.requested.kind, DeviceKind(r.requested.device)),
            ExecutionPolicy(r.resolved.kind, DeviceKind(r.resolved.device)),
            int(r.h2d_bytes), int(r.d2d_bytes), int(r.d2h_bytes),
            tuple(name for bit, name in _STAGE_NAMES if r.stage_mask & bit))


@dataclass
class ExecutionOutput:
    output: object
    receipt: ExecutionReceipt


@dataclass(frozen=True)
class Isometry3:
    """Rotation quaternion (x, y, z, w) and translation, f32 like Isometry3<f32>."""
    rotation: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 1.0)
    translation: Tuple[float, float, float] = (0.0, 0.0, 0.0)

    @staticmethod
    def identity():
        return Isometry3()

    def matrix(self) -> np.ndarray:
        x, y, z, w = (float(v) for v in self.rotation)
        n = np.sqrt(x * x + y * y + z * z + w * w)
        x, y, z, w = x / n, y / n, z / n, w / n
        m = np.eye(4, dtype=np.float32)
        m[:3, :3] = [[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]]
        m[:3, 3] = self.translation
        return m


@dataclass(frozen=True)
class GicpConfig:
    """GicpConfig with the reference defaults (gicp.rs:46-60) plus `voxel_leaf`
    (the centroid voxel filter applied to both clouds first; None = off)."""
    max_iterations: int = 50
    max_correspondence_distance: float = 1.0
    k_neighbors: int = 20
    epsilon: float = 1e-3
    transformation_epsilon: float = 1e-8
    fitness_epsilon: float = 1e-6
    min_correspondences: int = 6
    initial_guess: Isometry3 = field(default_factory=Isometry3.identity)
    covariance_radius: Optional[float] = None
    voxel_leaf: Optional[float] = None

    @staticmethod
    def with_correspondence_distance(d: float):  # gicp.rs:62-68
        return GicpConfig(max_correspondence_distance=d)

    def _c(self):
        c = _n.GicpConfigC()
        c.max_iterations = int(self.max_iterations)
        c.max_correspondence_distance = float(self.max_correspondence_distance)
        c.k_neighbors = int(self.k_neighbors)
        c.epsilon = float(self.epsilon)
        c.transformation_epsilon = float(self.transformation_epsilon)
        c.fitness_epsilon = float(self.fitness_epsilon)
        c.min_correspondences = int(self.min_correspondences)
        c.init_q[:] = [float(v) for v in self.initial_guess.rotation]
        c.init_t[:] = [float(v) for v in self.initial_guess.translation]
        c.covariance_radius = float(self.covariance_radius) if self.covariance_radius else 0.0
        c.voxel_leaf = float(self.voxel_leaf) if self.voxel_leaf else 0.0
        return c


@dataclass
class RegistrationResult:
    """RegistrationResult (registration.rs:15-24) plus the FP64 pose."""
    transform: Isometry3
    fitness: float
    iterations: int
    converged: bool
    rotation64: np.ndarray
    translation64: np.ndarray
    source_points: int
    target_points: int
    correspondences: int

    @staticmethod
    def _from_c(r):
        return RegistrationResult(
            Isometry3(tuple(float(v) for v in r.q), tuple(float(v) for v in r.t)),
            float(r.fitness), int(r.iterations), bool(r.converged),
            np.array(r.q64[:]), np.array(r.t64[:]), int(r.source_points), int(r.target_points),
            int(r.correspondences))

    def __repr__(self):
        return (f"RegistrationResult(fitness={self.fitness:.6g}, iterations={self.iterations}, "
                f"converged={self.converged})")


class _Cols:
    """Keeps contiguous f32 columns alive for a host sr_xyz_view."""

    def __init__(self, points):
        pts = np.asarray(points)
        if pts.ndim != 2 or pts.shape[1] != 3:
            raise InvalidArgument("points must have shape (N, 3)")
        self.x = np.ascontiguousarray(pts[:, 0], dtype=np.float32)
        self.y = np.ascontiguousarray(pts[:, 1], dtype=np.float32)
        self.z = np.ascontiguousarray(pts[:, 2], dtype=np.float32)
        self.view = _n.XyzView(self.x.ctypes.data, self.y.ctypes.data, self.z.ctypes.data,
                               len(self.x), _n.SR_MEM_HOST)


def device_view(x_ptr: int, y_ptr: int, z_ptr: int, n: int):
    """An sr_xyz_view over device-resident f32 columns (e.g. torch tensor data_ptr())."""
    return _n.XyzView(x_ptr, y_ptr, z_ptr, n, _n.SR_MEM_DEVICE)


def _check(code: int):
    if code == _n.SR_OK:
        return
    msg = _n.last_error()
    if code == _n.SR_STATUS_UNAVAILABLE:
        raise DeviceUnavailable(msg)
    raise InvalidArgument(msg)


def _view(points):
    if isinstance(points, _n.XyzView):
        return points, None
    cols = _Cols(points)
    return cols.view, cols


def device_available() -> bool:
    return bool(_n.load().sr_device_available())


def set_option(name: str, value: int) -> None:
    """Process-wide runtime option (sr_option_set), e.g. "match_reuse"."""
    _check(_n.load().sr_option_set(name.encode(), int(value)))


def get_option(name: str) -> int:
    v = ctypes.c_int64(0)
    _check(_n.load().sr_option_get(name.encode(), ctypes.byref(v)))
    return int(v.value)


def tune_describe() -> str:
    """The launch tuner's decisions so far (sr_tune_describe)."""
    n = ctypes.c_uint64(0)
    _check(_n.load().sr_tune_describe(None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(int(n.value) + 1)
    _check(_n.load().sr_tune_describe(buf, len(buf), ctypes.byref(n)))
    return buf.value.decode()


class GicpRegistration:
    """Generalized ICP on the HIP device (GicpRegistration, gicp.rs:70-235)."""

    def __init__(self, config: GicpConfig = GicpConfig()):
        self._config = config

    def config(self) -> GicpConfig:
        return self._config

    def name(self) -> str:
        return "GicpRegistration"

    def align(self, source, target) -> RegistrationResult:
        return self.align_with_policy(source, target, ExecutionPolicy.gpu())

    def align_with_diagnostics(self, source, target) -> RegistrationResult:
        return self.align(source, target)

    def align_with_policy(self, source, target, policy: ExecutionPolicy) -> RegistrationResult:
        return self.align_with_policy_and_receipt(source, target, policy).output

    def align_with_policy_and_receipt(self, source, target, policy: ExecutionPolicy):
        sv, _ks = _view(source)
        tv, _kt = _view(target)
        out = _n.ResultC()
        rc = _n.ReceiptC()
        cfg = self._config._c()
        _check(_n.load().sr_gicp_align(ctypes.byref(sv), ctypes.byref(tv), ctypes.byref(cfg),
                                       policy._c(), ctypes.byref(out), ctypes.byref(rc)))
        return ExecutionOutput(RegistrationResult._from_c(out), ExecutionReceipt._from_c(rc))


@dataclass(frozen=True)
class VoxelGridDownsampleConfig:
    leaf_size: float
    origin: Optional[Tuple[float, float, float]] = None

    @staticmethod
    def centroid(leaf_size: float):  # voxel.rs:138-146
        return VoxelGridDownsampleConfig(leaf_size)


@dataclass
class VoxelOutput:
    points: np.ndarray   # (C, 3) f32, first-occurrence order
    keys: np.ndarray     # (C, 3) int64
    counts: np.ndarray   # (C,) uint32


class VoxelGridDownsample:
    """Centroid voxel filter for xyz clouds (VoxelGridDownsample, voxel.rs:185-248)."""

    def __init__(self, config: VoxelGridDownsampleConfig):
        self._config = config

    def config(self):
        return self._config

    def filter(self, points) -> np.ndarray:
        return self.filter_with_policy(points, ExecutionPolicy.gpu()).points

    def filter_with_policy(self, points, policy: ExecutionPolicy) -> VoxelOutput:
        return self.filter_with_policy_and_receipt(points, policy).output

    def filter_with_policy_and_receipt(self, points, policy: ExecutionPolicy):
        pts = np.asarray(points, dtype=np.float32).reshape(-1, 3)
        n = len(pts)
        cols = _Cols(pts) if n else None
        view = cols.view if cols else _n.XyzView(None, None, None, 0, _n.SR_MEM_HOST)
        cap = max(n, 1)
        ox, oy, oz = (np.zeros(cap, np.float32) for _ in range(3))
        keys = np.zeros((cap, 3), np.int64)
        counts = np.zeros(cap, np.uint32)
        n_out = ctypes.c_uint64(0)
        rc = _n.ReceiptC()
        org = None if self._config.origin is None else np.asarray(self._config.origin, np.float32)
        _check(_n.load().sr_voxel_downsample(
            ctypes.byref(view), float(self._config.leaf_size),
            None if org is None else org.ctypes.data, policy._c(), ox.ctypes.data, oy.ctypes.data,
            oz.ctypes.data, keys.ctypes.data, counts.ctypes.data, ctypes.byref(n_out),
            ctypes.byref(rc)))
        c = n_out.value
        out = VoxelOutput(np.stack([ox[:c], oy[:c], oz[:c]], 1), keys[:c].copy(), counts[:c].copy())
        return ExecutionOutput(out, ExecutionReceipt._from_c(rc))


def plane_covariances(points, k_neighbors: int = 20, radius: Optional[float] = None,
                      epsilon: float = 1e-3):
    """Per-point regularized covariances [c00,c11,c22,c01,c02,c12] (N,6) and unit
    normals (N,3): k-NN (gicp.rs:259-291) or radius grid (covariances_grid.rs:216)."""
    cols = _Cols(points)
    n = len(cols.x)
    cov = np.zeros((max(n, 1), 6), np.float32)
    nrm = np.zeros((max(n, 1), 3), np.float32)
    _check(_n.load().sr_plane_covariances(ctypes.byref(cols.view), int(k_neighbors),
                                          float(radius) if radius else 0.0, float(epsilon),
                                          cov.ctypes.data, nrm.ctypes.data))
    return cov[:n], nrm[:n]


@dataclass
class NormalEstimationConfig:
    """The radius-mode fields of NormalEstimationConfig (features/src/normal.rs)."""
    search_radius: float = 0.1
    viewpoint: Optional[Tuple[float, float, float]] = None


class GpuNormalEstimator:
    """GpuNormalEstimator with a search radius (features/src/normal_gpu.rs:33-82) on
    the HIP device: estimate() returns (normals (N,3) f32, curvature (N,) f32)."""

    def __init__(self, config: NormalEstimationConfig):
        self._config = config

    def config(self) -> NormalEstimationConfig:
        return self._config

    def estimate(self, points):
        cols = _Cols(points)
        n = len(cols.x)
        nrm = np.zeros((max(n, 1), 3), np.float32)
        curv = np.zeros(max(n, 1), np.float32)
        vp = self._config.viewpoint
        vpa = None if vp is None else np.ascontiguousarray(vp, dtype=np.float32)
        _check(_n.load().sr_grid_normals(ctypes.byref(cols.view), float(self._config.search_radius),
                                         vpa.ctypes.data if vpa is not None else None,
                                         nrm.ctypes.data, curv.ctypes.data))
        return nrm[:n], curv[:n]


def knn(points, queries, k: int):
    """Exact k-NN (ascending distance, ties to the lower index); -1 pads."""
    pc = _Cols(points)
    qc = _Cols(queries)
    nq = len(qc.x)
    idx = np.zeros((max(nq, 1), max(k, 1)), np.int32)
    d2 = np.zeros((max(nq, 1), max(k, 1)), np.float32)
    _check(_n.load().sr_knn(ctypes.byref(pc.view), ctypes.byref(qc.view), int(k), idx.ctypes.data,
                            d2.ctypes.data))
    return idx[:nq, :k], d2[:nq, :k]


def gicp_normal_equations(source, target, config: "GicpConfig", q, t):
    """One device correspondence-and-accumulate pass at the FP64 pose (q, t)
    (gicp.rs:137-179, no solve): (match (N,) int32 with -1 for none,
    H (6,6), g (6,), count, sum of squared distances)."""
    sv, keep_s = _view(source)
    tv, keep_t = _view(target)
    n = int(sv.n)
    match = np.zeros(max(n, 1), np.int32)
    H = np.zeros((6, 6))
    g = np.zeros(6)
    cnt = ctypes.c_uint64(0)
    d2 = ctypes.c_double(0.0)
    qa = np.ascontiguousarray(q, dtype=np.float64)
    ta = np.ascontiguousarray(t, dtype=np.float64)
    _check(_n.load().sr_gicp_normal_equations(
        ctypes.byref(sv), ctypes.byref(tv), ctypes.byref(config._c()), qa.ctypes.data,
        ta.ctypes.data, match.ctypes.data, H.ctypes.data, g.ctypes.data, ctypes.addressof(cnt),
        ctypes.addressof(d2)))
    del keep_s, keep_t
    return match[:n], H, g, int(cnt.value), float(d2.value)


class TargetMap:
    """A target cloud preprocessed once and kept on the device (sr_map)."""

    def __init__(self, target, config: GicpConfig):
        tv, self._keep = _view(target)
        self._cfg = config
        h = ctypes.c_void_p()
        _check(_n.load().sr_map_create(ctypes.byref(tv), ctypes.byref(config._c()), ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def points(self) -> int:
        return int(_n.load().sr_map_points(self._h))

    def device_bytes(self) -> int:
        return int(_n.load().sr_map_device_bytes(self._h))

    def align(self, source, config: Optional[GicpConfig] = None) -> RegistrationResult:
        sv, _k = _view(source)
        out = _n.ResultC()
        cfg = (config or self._cfg)._c()
        _check(_n.load().sr_map_align(self._h, ctypes.byref(sv), ctypes.byref(cfg), ctypes.byref(out)))
        return RegistrationResult._from_c(out)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _n.load().sr_map_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_align(maps, sources, config: GicpConfig):
    """Registers every (source, map) pair in one set of launches."""
    count = len(maps)
    keep = []
    views = (_n.XyzView * max(count, 1))()
    for i, s in enumerate(sources):
        v, k = _view(s)
        views[i] = v
        keep.append(k)
    handles = (ctypes.c_void_p * max(count, 1))(*[m.handle.value for m in maps])
    out = (_n.ResultC * max(count, 1))()
    _check(_n.load().sr_batch_align(handles, views, count, ctypes.byref(config._c()), out))
    return [RegistrationResult._from_c(out[i]) for i in range(count)]


class GicpScanMatcher:
    """ScanMatcher (crates/spatialrust-mapping/src/scan.rs:12-20) backed by GICP
    on the device: match_scans(previous, current) estimates current_T_previous,
    i.e. registers `previous` (source) onto `current` (target), as
    IcpScanMatcher does with ICP (scan.rs:219-250)."""

    def __init__(self, config: GicpConfig):
        self._reg = GicpRegistration(config)

    def config(self) -> GicpConfig:
        return self._reg.config()

    def match_scans(self, previous, current) -> Isometry3:
        return self._reg.align(previous, current).transform


class RollingMapOdometry:
    """Scan-to-map odometry over a rolling device-resident voxel map
    (sr_odometry): push() registers a sensor-frame scan against the map from a
    constant-velocity guess, merges its new voxels and returns map_T_scan."""

    def __init__(self, config: GicpConfig, voxel_leaf: float = 0.25,
                 max_map_points: int = 4_000_000):
        oc = _n.OdometryConfigC(float(voxel_leaf), 0, int(max_map_points))
        h = ctypes.c_void_p()
        _check(_n.load().sr_odometry_create(ctypes.byref(oc), ctypes.byref(config._c()),
                                            ctypes.byref(h)))
        self._h = h

    def push(self, scan) -> RegistrationResult:
        sv, _k = _view(scan)
        out = _n.ResultC()
        _check(_n.load().sr_odometry_push(self._h, ctypes.byref(sv), ctypes.byref(out)))
        return RegistrationResult._from_c(out)

    def map_points(self) -> int:
        return int(_n.load().sr_odometry_map_points(self._h))

    @staticmethod
    def push_many(odometries, scans) -> list:
        """One scan for each of several distinct odometries, registered in one
        batch (sr_odometry_push_many); the results equal push() on each."""
        if len(odometries) != len(scans):
            raise InvalidArgument("one scan per odometry")
        n = len(odometries)
        if n == 0:
            return []
        views = [_view(s) for s in scans]
        arr = (_n.XyzView * n)(*[v for v, _ in views])
        hs = (ctypes.c_void_p * n)(*[o._h.value for o in odometries])
        out = (_n.ResultC * n)()
        _check(_n.load().sr_odometry_push_many(hs, arr, n, out))
        return [RegistrationResult._from_c(out[i]) for i in range(n)]

    def map_stats(self) -> dict:
        """Map upkeep counters: map points, normals computed by the last
        push, full grid rebuilds so far, scans pushed."""
        out = (ctypes.c_uint64 * 4)()
        _check(_n.load().sr_odometry_map_stats(self._h, out))
        return {"map_points": int(out[0]), "refreshed": int(out[1]),
                "rebuilds": int(out[2]), "scans": int(out[3])}

    def map_cloud(self) -> np.ndarray:
        n = self.map_points()
        out = np.zeros((max(n, 1), 3), np.float32)
        got = ctypes.c_uint64(0)
        _check(_n.load().sr_odometry_map_read(self._h, out.ctypes.data, n, ctypes.byref(got)))
        return out[:n]

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _n.load().sr_odometry_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching registration server over resident maps (sr_server):
    submit() queues a source, step() runs one scheduling step (admission by
    slots and HBM budget, one Gauss-Newton pass over every active request,
    retirement), poll() fetches a finished result."""

    def __init__(self, config: GicpConfig, hbm_budget_bytes: int = 200 << 30,
                 max_active: int = 256, max_iterations: int = 0,
                 delta_pose_epsilon: float = 0.0):
        sc = _n.ServerConfigC(int(hbm_budget_bytes), int(max_active), int(max_iterations),
                              float(delta_pose_epsilon))
        h = ctypes.c_void_p()
        _check(_n.load().sr_server_create(ctypes.byref(sc), ctypes.byref(config._c()),
                                          ctypes.byref(h)))
        self._h = h

    def submit(self, target_map: "TargetMap", source) -> int:
        sv, _k = _view(source)
        t = ctypes.c_uint64(0)
        _check(_n.load().sr_server_submit(self._h, target_map.handle, ctypes.byref(sv),
                                          ctypes.byref(t)))
        return int(t.value)

    def step(self) -> int:
        pending = ctypes.c_uint32(0)
        _check(_n.load().sr_server_step(self._h, ctypes.byref(pending)))
        return int(pending.value)

    def latency(self, ticket: int) -> float:
        sec = ctypes.c_double(0.0)
        _check(_n.load().sr_server_latency(self._h, int(ticket), ctypes.byref(sec)))
        return float(sec.value)

    def poll(self, ticket: int) -> Optional[RegistrationResult]:
        out = _n.ResultC()
        ready = ctypes.c_uint8(0)
        _check(_n.load().sr_server_poll(self._h, int(ticket), ctypes.byref(out), ctypes.byref(ready)))
        return RegistrationResult._from_c(out) if ready.value else None

    def device_bytes(self) -> int:
        return int(_n.load().sr_server_device_bytes(self._h))

    def finished(self, cap: int = 4096):
        """Tickets finished since the last call (at most `cap`)."""
        buf = (ctypes.c_uint64 * max(cap, 1))()
        n = ctypes.c_uint64(0)
        _check(_n.load().sr_server_finished(self._h, buf, cap, ctypes.byref(n)))
        return [int(buf[i]) for i in range(n.value)]

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _n.load().sr_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_align_pairs(sources, targets, config: GicpConfig):
    """Registers every (source, target) pair: all clouds preprocessed on the
    device's worker streams, then one set of Gauss-Newton launches."""
    count = len(sources)
    keep = []
    sv = (_n.XyzView * max(count, 1))()
    tv = (_n.XyzView * max(count, 1))()
    for i, (a, b) in enumerate(zip(sources, targets)):
        va, ka = _view(a)
        vb, kb = _view(b)
        sv[i], tv[i] = va, vb
        keep += [ka, kb]
    out = (_n.ResultC * max(count, 1))()
    _check(_n.load().sr_batch_align_pairs(sv, tv, count, ctypes.byref(config._c()), out))
    return [RegistrationResult._from_c(out[i]) for i in range(count)]


def register_gicp(source, target, max_correspondence_distance: float = 1.0,
                  max_iterations: int = 50, k_neighbors: int = 20) -> RegistrationResult:
    """register_gicp (crates/spatialrust-py/src/lib.rs:2103-2121), same signature and defaults."""
    cfg = GicpConfig(max_iterations=max_iterations,
                     max_correspondence_distance=max_correspondence_distance,
                     k_neighbors=k_neighbors)
    return GicpRegistration(cfg).align(source, target)
