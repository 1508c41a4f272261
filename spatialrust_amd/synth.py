"""Seeded synthetic scenes for the registration workloads (numpy, f32 output).

Generator (the reference ships none; this one is fixed so every run sees the
same clouds):
  - RNG: counter-based SplitMix64 -> uniform doubles in [0, 1).
  - Gaussian noise: Box-Muller in f64, cast to f32 after adding.
  - Surfaces are sampled area-uniformly: a primitive is chosen with probability
    proportional to its area, then a point uniformly on it.
  - Scenes are centred on the origin in x/y so a rigid offset rotates about the
    scene centre.

Workloads
  small_cube    20 m cube, ground + 4 walls, 12 boxes, 6 spheres, sigma 0.01
  large_block   100 x 100 x 10 m, ground + 4 walls, 300 boxes, 100 cylinders, sigma 0.02
"""
from dataclasses import dataclass, field

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix_uniform(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n uniform doubles in [0,1) from SplitMix64 at counters 1..n of (seed, stream)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF)
        z = base + (np.arange(1, n + 1, dtype=np.uint64) * _GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class Rng:
    """Sequential draws from independent SplitMix64 streams of one seed."""

    def __init__(self, seed: int):
        self.seed = int(seed)
        self.stream = 0

    def uniform(self, n, lo=0.0, hi=1.0):
        self.stream += 1
        return lo + (hi - lo) * splitmix_uniform(self.seed, int(n), self.stream)

    def normal(self, n):
        u1 = np.maximum(self.uniform(n), 1e-300)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)

    def choice(self, n, weights):
        cdf = np.cumsum(weights) / np.sum(weights)
        return np.minimum(np.searchsorted(cdf, self.uniform(n), side="right"), len(weights) - 1)


@dataclass
class Scene:
    """Axis-aligned rectangles, spheres and vertical cylinders (lateral + top)."""
    rects: list = field(default_factory=list)      # (origin[3], u[3], v[3])
    boxes: list = field(default_factory=list)      # (lo[3], hi[3]) of box primitives (their faces are in rects)
    spheres: list = field(default_factory=list)    # (center[3], r)
    cylinders: list = field(default_factory=list)  # (cx, cy, z0, r, h)
    extent: float = 1.0

    def areas(self):
        a = [float(np.linalg.norm(np.cross(u, v))) for (_, u, v) in self.rects]
        a += [4.0 * np.pi * r * r for (_, r) in self.spheres]
        a += [2.0 * np.pi * r * h + np.pi * r * r for (_, _, _, r, h) in self.cylinders]
        return np.array(a, dtype=np.float64)


def _box_faces(lo, hi, with_bottom=False):
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    d = hi - lo
    ex, ey, ez = np.array([d[0], 0, 0]), np.array([0, d[1], 0]), np.array([0, 0, d[2]])
    faces = [
        (lo + np.array([0, 0, d[2]]), ex, ey),  # top
        (lo, ex, ez), (lo + np.array([0, d[1], 0]), ex, ez),  # y faces
        (lo, ey, ez), (lo + np.array([d[0], 0, 0]), ey, ez),  # x faces
    ]
    if with_bottom:
        faces.append((lo, ex, ey))
    return faces


def small_cube_scene(seed: int) -> Scene:
    """20 m cube centred on the origin (x,y in [-10,10], z in [0,20]): ground,
    four walls, 12 boxes (edges U(0.5,3) m) on the ground, 6 spheres (r U(0.3,1.5))."""
    rng = Rng(seed)
    s = Scene(extent=20.0)
    L = 20.0
    h = L / 2
    s.rects.append((np.array([-h, -h, 0.0]), np.array([L, 0, 0.0]), np.array([0, L, 0.0])))
    s.rects.append((np.array([-h, -h, 0.0]), np.array([L, 0, 0.0]), np.array([0, 0, L])))
    s.rects.append((np.array([-h, h, 0.0]), np.array([L, 0, 0.0]), np.array([0, 0, L])))
    s.rects.append((np.array([-h, -h, 0.0]), np.array([0, L, 0.0]), np.array([0, 0, L])))
    s.rects.append((np.array([h, -h, 0.0]), np.array([0, L, 0.0]), np.array([0, 0, L])))
    e = rng.uniform(36, 0.5, 3.0).reshape(12, 3)
    c = rng.uniform(24, -h + 2.0, h - 2.0).reshape(12, 2)
    for i in range(12):
        lo = np.array([c[i, 0] - e[i, 0] / 2, c[i, 1] - e[i, 1] / 2, 0.0])
        s.rects.extend(_box_faces(lo, lo + e[i]))
        s.boxes.append((lo, lo + e[i]))
    r = rng.uniform(6, 0.3, 1.5)
    cc = rng.uniform(12, -h + 2.0, h - 2.0).reshape(6, 2)
    cz = rng.uniform(6, 0.0, 4.0)
    for i in range(6):
        s.spheres.append((np.array([cc[i, 0], cc[i, 1], r[i] + cz[i]]), float(r[i])))
    return s


def large_block_scene(seed: int) -> Scene:
    """100 x 100 x 10 m block centred on the origin in x/y: ground, four walls,
    300 boxes (footprint U(0.5,6) m, height U(0.5,10) m), 100 cylinders
    (r U(0.2,1.5), height U(1,10))."""
    rng = Rng(seed)
    s = Scene(extent=100.0)
    L, H = 100.0, 10.0
    h = L / 2
    s.rects.append((np.array([-h, -h, 0.0]), np.array([L, 0, 0.0]), np.array([0, L, 0.0])))
    s.rects.append((np.array([-h, -h, 0.0]), np.array([L, 0, 0.0]), np.array([0, 0, H])))
    s.rects.append((np.array([-h, h, 0.0]), np.array([L, 0, 0.0]), np.array([0, 0, H])))
    s.rects.append((np.array([-h, -h, 0.0]), np.array([0, L, 0.0]), np.array([0, 0, H])))
    s.rects.append((np.array([h, -h, 0.0]), np.array([0, L, 0.0]), np.array([0, 0, H])))
    fp = rng.uniform(600, 0.5, 6.0).reshape(300, 2)
    ht = rng.uniform(300, 0.5, H)
    c = rng.uniform(600, -h + 4.0, h - 4.0).reshape(300, 2)
    for i in range(300):
        lo = np.array([c[i, 0] - fp[i, 0] / 2, c[i, 1] - fp[i, 1] / 2, 0.0])
        s.rects.extend(_box_faces(lo, lo + np.array([fp[i, 0], fp[i, 1], ht[i]])))
        s.boxes.append((lo, lo + np.array([fp[i, 0], fp[i, 1], ht[i]])))
    r = rng.uniform(100, 0.2, 1.5)
    hc = rng.uniform(100, 1.0, H)
    cc = rng.uniform(200, -h + 3.0, h - 3.0).reshape(100, 2)
    for i in range(100):
        s.cylinders.append((float(cc[i, 0]), float(cc[i, 1]), 0.0, float(r[i]), float(hc[i])))
    return s


def sample_scene(scene: Scene, n: int, sigma: float, seed: int) -> np.ndarray:
    """n points area-uniform on the scene surfaces plus isotropic Gaussian noise
    (sigma, metres). Returns (n, 3) float32."""
    rng = Rng(seed)
    areas = scene.areas()
    prim = rng.choice(n, areas)
    a = rng.uniform(n)
    b = rng.uniform(n)
    out = np.zeros((n, 3), np.float64)
    nr, ns = len(scene.rects), len(scene.spheres)
    if nr:
        o = np.array([r[0] for r in scene.rects])
        u = np.array([r[1] for r in scene.rects])
        v = np.array([r[2] for r in scene.rects])
        m = prim < nr
        k = prim[m]
        out[m] = o[k] + a[m, None] * u[k] + b[m, None] * v[k]
    if ns:
        m = (prim >= nr) & (prim < nr + ns)
        k = prim[m] - nr
        cen = np.array([sp[0] for sp in scene.spheres])
        rad = np.array([sp[1] for sp in scene.spheres])
        zc = 2.0 * a[m] - 1.0
        ph = 2.0 * np.pi * b[m]
        rr = np.sqrt(np.maximum(0.0, 1.0 - zc * zc))
        out[m] = cen[k] + rad[k, None] * np.stack([rr * np.cos(ph), rr * np.sin(ph), zc], 1)
    if scene.cylinders:
        m = prim >= nr + ns
        k = prim[m] - nr - ns
        cyl = np.array(scene.cylinders)
        cx, cy, z0, r, h = (cyl[k, j] for j in range(5))
        lateral = 2.0 * np.pi * r * h
        top = np.pi * r * r
        on_top = a[m] * (lateral + top) > lateral
        ang = 2.0 * np.pi * b[m]
        # lateral: height from the rescaled first draw
        t_lat = np.where(on_top, 0.0, a[m] * (lateral + top) / np.maximum(lateral, 1e-12))
        px = cx + r * np.cos(ang)
        py = cy + r * np.sin(ang)
        pz = z0 + t_lat * h
        # top disk: radius sqrt of the rescaled draw
        rho = r * np.sqrt(np.clip((a[m] * (lateral + top) - lateral) / np.maximum(top, 1e-12), 0, 1))
        tx = cx + rho * np.cos(ang)
        ty = cy + rho * np.sin(ang)
        out[m] = np.where(on_top[:, None], np.stack([tx, ty, z0 + h], 1), np.stack([px, py, pz], 1))
    noise = np.stack([rng.normal(n), rng.normal(n), rng.normal(n)], 1) * sigma
    return (out + noise).astype(np.float32)


# ---------------------------------------------------------- spinning LiDAR --

ODOMETRY_RING = 30.0      # loop radius of the odometry trajectory (m)
ODOMETRY_CORRIDOR = 4.0   # obstacles within this distance of the loop are removed


def odometry_scene(seed: int) -> Scene:
    """The large_block scene with the boxes and cylinders that reach within
    ODOMETRY_CORRIDOR of the odometry loop (a circle of radius ODOMETRY_RING about
    the origin) removed, so the sensor drives through free space."""
    src = large_block_scene(seed)
    s = Scene(extent=src.extent)
    s.rects = list(src.rects[:5])  # ground and walls
    lo_r, hi_r = ODOMETRY_RING - ODOMETRY_CORRIDOR, ODOMETRY_RING + ODOMETRY_CORRIDOR
    for lo, hi in src.boxes:
        nx = np.clip(0.0, lo[0], hi[0])
        ny = np.clip(0.0, lo[1], hi[1])
        dmin = np.hypot(nx, ny)
        dmax = max(np.hypot(x, y) for x in (lo[0], hi[0]) for y in (lo[1], hi[1]))
        if dmin <= hi_r and dmax >= lo_r:
            continue
        s.boxes.append((lo, hi))
        s.rects.extend(_box_faces(lo, hi))
    for (cx, cy, z0, r, h) in src.cylinders:
        d = np.hypot(cx, cy)
        if d - r <= hi_r and d + r >= lo_r:
            continue
        s.cylinders.append((cx, cy, z0, r, h))
    return s


def odometry_trajectory(n_scans: int, seed: int, speed: float = 10.0, hz: float = 10.0,
                        height: float = 1.8):
    """world_T_sensor poses (q (x,y,z,w), t) of a smooth loop at `speed` m/s
    sampled at `hz`: radius ODOMETRY_RING + 1.5 sin(2 phi + seeded phase), heading
    along the path (yaw rate about 20 deg/s, below 30 deg/s)."""
    ph = float(Rng(seed).uniform(1, 0.0, 2.0 * np.pi)[0])
    dphi = 1e-3
    phis = [0.0]
    # arc-length parametrisation of r(phi) by fine integration
    def radius(phi):
        return ODOMETRY_RING + 1.5 * np.sin(2.0 * phi + ph)
    step = speed / hz
    phi, acc = 0.0, 0.0
    while len(phis) < n_scans:
        r0, r1 = radius(phi), radius(phi + dphi)
        p0 = np.array([r0 * np.cos(phi), r0 * np.sin(phi)])
        p1 = np.array([r1 * np.cos(phi + dphi), r1 * np.sin(phi + dphi)])
        acc += float(np.linalg.norm(p1 - p0))
        phi += dphi
        if acc >= step:
            phis.append(phi)
            acc -= step
    poses = []
    for phi in phis:
        r = radius(phi)
        p = np.array([r * np.cos(phi), r * np.sin(phi)])
        r2 = radius(phi + dphi)
        p2 = np.array([r2 * np.cos(phi + dphi), r2 * np.sin(phi + dphi)])
        yaw = float(np.arctan2(p2[1] - p[1], p2[0] - p[0]))
        q = np.array([0.0, 0.0, np.sin(yaw / 2.0), np.cos(yaw / 2.0)])
        poses.append((q, np.array([p[0], p[1], height])))
    return poses


def lidar_scan(scene: Scene, q, t, seed: int, beams: int = 64, azimuth_bins: int = 2048,
               elevation_deg=(-25.0, 15.0), sigma: float = 0.02, dropout: float = 0.05,
               max_range: float = 120.0, device: str = "cpu") -> np.ndarray:
    """One spinning-LiDAR scan from sensor pose (q, t) = world_T_sensor: beams
    with elevations linearly spaced over elevation_deg, azimuth_bins per
    revolution, ray-cast against the scene's ground and walls (rects[:5]), boxes
    (slab test) and cylinders (lateral surface and top). Range noise N(0, sigma)
    and a random `dropout` fraction removed. Returns sensor-frame xyz (N, 3) f32.
    The ray casting runs in FP64 torch on `device` (the result does not depend
    on it beyond FP64 rounding of the hit distance)."""
    import torch
    rng = Rng(seed)
    el = np.deg2rad(np.linspace(elevation_deg[0], elevation_deg[1], beams))
    az = 2.0 * np.pi * np.arange(azimuth_bins) / azimuth_bins
    E, A = np.meshgrid(el, az, indexing="ij")
    d_s = np.stack([np.cos(E) * np.cos(A), np.cos(E) * np.sin(A), np.sin(E)], -1).reshape(-1, 3)
    R = quat_to_mat(q)
    f64 = dict(dtype=torch.float64, device=device)
    D = torch.as_tensor(d_s @ R.T, **f64)
    O = torch.as_tensor(np.asarray(t, np.float64), **f64)
    inf = torch.tensor(float("inf"), **f64)
    best = torch.full((D.shape[0],), float("inf"), **f64)
    inv = 1.0 / D
    # ground and walls: axis-aligned rectangles
    for (o, u, v) in scene.rects[:5]:
        o, u, v = (np.asarray(a, np.float64) for a in (o, u, v))
        ax = int(np.nonzero((u == 0) & (v == 0))[0][0])
        tt = (float(o[ax]) - O[ax]) * inv[:, ax]
        ok = tt > 1e-6
        for b in range(3):
            if b == ax:
                continue
            lo_b, hi_b = sorted((float(o[b]), float(o[b] + u[b] + v[b])))
            hb = O[b] + tt * D[:, b]
            ok &= (hb >= lo_b) & (hb <= hi_b)
        best = torch.where(ok & (tt < best), tt, best)
    # boxes: slab test
    if scene.boxes:
        B_lo = torch.as_tensor(np.array([b[0] for b in scene.boxes]), **f64)
        B_hi = torch.as_tensor(np.array([b[1] for b in scene.boxes]), **f64)
        for c0 in range(0, B_lo.shape[0], 16):
            t1 = (B_lo[None, c0:c0 + 16] - O) * inv[:, None, :]
            t2 = (B_hi[None, c0:c0 + 16] - O) * inv[:, None, :]
            tmin = torch.minimum(t1, t2).nan_to_num(nan=-float("inf")).amax(dim=2)
            tmax = torch.maximum(t1, t2).nan_to_num(nan=float("inf")).amin(dim=2)
            ok = (tmax >= tmin) & (tmin > 1e-6)
            best = torch.minimum(best, torch.where(ok, tmin, inf).amin(dim=1))
    # cylinders: lateral surface (entry root) and top disc
    if scene.cylinders:
        C = torch.as_tensor(np.array(scene.cylinders, np.float64), **f64)  # (M, 5)
        for c0 in range(0, C.shape[0], 16):
            cx, cy, z0, r, h = (C[c0:c0 + 16, k][None] for k in range(5))
            ox, oy = O[0] - cx, O[1] - cy
            dx, dy, dz = D[:, 0:1], D[:, 1:2], D[:, 2:3]
            a = dx * dx + dy * dy
            bb = ox * dx + oy * dy
            cc = ox * ox + oy * oy - r * r
            disc = bb * bb - a * cc
            tl = (-bb - torch.sqrt(torch.clamp(disc, min=0.0))) / a
            zl = O[2] + tl * dz
            okl = (disc >= 0) & (tl > 1e-6) & (zl >= z0) & (zl <= z0 + h)
            tt = (z0 + h - O[2]) / dz
            hx = ox + tt * dx
            hy = oy + tt * dy
            okt = (tt > 1e-6) & (hx * hx + hy * hy <= r * r)
            cand = torch.minimum(torch.where(okl, tl, inf), torch.where(okt, tt, inf)).amin(dim=1)
            best = torch.minimum(best, cand)
    best = best.cpu().numpy()
    keep = np.isfinite(best) & (best <= max_range)
    rng_noisy = best + sigma * rng.normal(len(best))
    keep &= rng.uniform(len(best)) >= dropout
    pts = d_s[keep] * rng_noisy[keep][:, None]
    return pts.astype(np.float32)


# ------------------------------------------------------------ rigid motion --

def odometry_scans(n: int, trajectory_scans: int = 0, device: str = "", progress: bool = False):
    """The first n full-size scans (64 x 2048 beams, seeds 100 + k) of the
    odometry workload along odometry_trajectory(max(n, trajectory_scans),
    seed=1) through odometry_scene(SEED_BASE + 4000). The ray casting runs on
    `device` ("" picks cuda when torch sees a GPU, else cpu). torch must own
    the process's first HIP initialisation: run this in a fresh process (e.g.
    a spawned worker) when another HIP library is already loaded."""
    import sys
    import torch
    dev = device or ("cuda" if torch.cuda.is_available() else "cpu")
    if dev == "cpu":
        torch.set_num_threads(min(8, torch.get_num_threads()))
    scene = odometry_scene(SEED_BASE + 4000)
    poses = odometry_trajectory(max(n, trajectory_scans), seed=1)[:n]
    out = []
    for k, (q, t) in enumerate(poses):
        out.append(lidar_scan(scene, q, t, seed=100 + k, device=dev))
        if progress and (k + 1) % 25 == 0:
            print(f"  scans ray-cast: {k + 1}/{n} ({dev})", file=sys.stderr, flush=True)
    return out


def quat_axis_angle(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    s = np.sin(angle / 2.0)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, np.cos(angle / 2.0)])


def quat_to_mat(q):
    """Rotation matrix of a unit quaternion (x, y, z, w), float64."""
    x, y, z, w = (float(v) for v in q)
    n = np.sqrt(x * x + y * y + z * z + w * w)
    x, y, z, w = x / n, y / n, z / n, w / n
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def quat_mul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw,
                     aw * bw - ax * bx - ay * by - az * bz])


def random_motion(seed: int, max_angle_deg: float, max_trans: float, per_axis: bool):
    """Rotation with a uniform axis and angle U(0, max_angle); translation
    U(-max_trans, max_trans) per axis (per_axis) or a uniform direction scaled by
    U(0, max_trans). Returns (q (x,y,z,w) float64, t float64)."""
    rng = Rng(seed ^ 0x5A5A5A5A)
    zc = rng.uniform(1, -1.0, 1.0)[0]
    ph = rng.uniform(1, 0.0, 2.0 * np.pi)[0]
    rr = np.sqrt(max(0.0, 1.0 - zc * zc))
    axis = np.array([rr * np.cos(ph), rr * np.sin(ph), zc])
    angle = np.deg2rad(rng.uniform(1, 0.0, max_angle_deg)[0])
    q = quat_axis_angle(axis, angle)
    if per_axis:
        t = rng.uniform(3, -max_trans, max_trans)
    else:
        zc2 = rng.uniform(1, -1.0, 1.0)[0]
        ph2 = rng.uniform(1, 0.0, 2.0 * np.pi)[0]
        r2 = np.sqrt(max(0.0, 1.0 - zc2 * zc2))
        t = np.array([r2 * np.cos(ph2), r2 * np.sin(ph2), zc2]) * rng.uniform(1, 0.0, max_trans)[0]
    return q, t


def apply_motion(q, t, pts):
    """p' = R p + t in float64, rounded to float32."""
    R = quat_to_mat(q)
    return (np.asarray(pts, np.float64) @ R.T + np.asarray(t, np.float64)).astype(np.float32)


def invert_motion(q, t):
    qi = np.array([-q[0], -q[1], -q[2], q[3]])
    ti = -(quat_to_mat(qi) @ np.asarray(t, np.float64))
    return qi, ti


def rotation_angle_between(qa, qb):
    """Angle (rad) of the relative rotation between two unit quaternions:
    2 atan2(|v|, |w|) of conj(qa) * qb = (v, w), which stays accurate for tiny
    angles (an arccos of the dot product rounds everything below ~2e-8 rad to 0)."""
    qa = np.asarray(qa, np.float64) / np.linalg.norm(qa)
    qb = np.asarray(qb, np.float64) / np.linalg.norm(qb)
    rel = quat_mul(np.array([-qa[0], -qa[1], -qa[2], qa[3]]), qb)
    return float(2.0 * np.arctan2(np.linalg.norm(rel[:3]), abs(rel[3])))


# --------------------------------------------------------------- workloads --

SEED_BASE = 0x5EED0000


@dataclass
class Workload:
    name: str
    source: np.ndarray
    target: np.ndarray
    gt_q: np.ndarray   # source -> target ground truth (inverse of the applied motion)
    gt_t: np.ndarray
    voxel_leaf: float
    k_neighbors: int
    max_correspondence_distance: float
    max_iterations: int
    extent: float


def small_cube_pair(instance: int = 0, n_source: int = 50_000, n_target: int = 50_000,
                    max_angle_deg: float = 10.0, max_trans: float = 0.5) -> Workload:
    """50k + 50k points of the 20 m cube; source moved by a seeded motion
    (angle U(0,10 deg), translation U(-0.5,0.5) per axis); leaf 0.1, k 20,
    d_max 1.0, 30 iterations."""
    seed = SEED_BASE + 1000 * 0 + instance
    scene = small_cube_scene(seed)
    tgt = sample_scene(scene, n_target, 0.01, seed * 7 + 1)
    src0 = sample_scene(scene, n_source, 0.01, seed * 7 + 2)
    q, t = random_motion(seed, max_angle_deg, max_trans, per_axis=True)
    src = apply_motion(q, t, src0)
    gq, gt = invert_motion(q, t)
    return Workload("small_cube_50k", src, tgt, gq, gt, 0.1, 20, 1.0, 30, 20.0)


def large_block_pair(n_source: int = 200_000, n_target: int = 2_000_000,
                     max_angle_deg: float = 15.0, max_trans: float = 2.0) -> Workload:
    """2M-point target and 200k-point source of the 100 x 100 x 10 m block;
    source offset by up to 15 deg / 2 m; leaf 0.2, k 15, d_max 2.0, 50 iterations."""
    seed = SEED_BASE + 1000 * 1
    scene = large_block_scene(seed)
    tgt = sample_scene(scene, n_target, 0.02, seed * 7 + 1)
    src0 = sample_scene(scene, n_source, 0.02, seed * 7 + 2)
    q, t = random_motion(seed, max_angle_deg, max_trans, per_axis=False)
    src = apply_motion(q, t, src0)
    gq, gt = invert_motion(q, t)
    return Workload("large_block_2M_200k", src, tgt, gq, gt, 0.2, 15, 2.0, 50, 100.0)


def batch_pairs(count: int = 256, lo: int = 5_000, hi: int = 120_000, seed_offset: int = 0):
    """Independent instances of the small-cube generator: source sizes
    log-uniform in [lo, hi], target 1.5x the source, offsets up to 10 deg / 1 m."""
    rng = Rng(SEED_BASE + 1000 * 2 + seed_offset)
    sizes = np.exp(rng.uniform(count, np.log(lo), np.log(hi))).astype(np.int64)
    out = []
    for i in range(count):
        seed = SEED_BASE + 1000 * 2 + 17 * i + seed_offset
        scene = small_cube_scene(seed)
        ns = int(sizes[i])
        nt = int(sizes[i] * 3 // 2)
        tgt = sample_scene(scene, nt, 0.01, seed * 7 + 1)
        src0 = sample_scene(scene, ns, 0.01, seed * 7 + 2)
        q, t = random_motion(seed, 10.0, 1.0, per_axis=False)
        gq, gt = invert_motion(q, t)
        out.append(Workload(f"batch_{i}", apply_motion(q, t, src0), tgt, gq, gt, 0.1, 20, 1.0, 30,
                            20.0))
    return out


# ------------------------------------------------------------------ serving --

@dataclass
class ServingPool:
    """Maps and request sources of the serving workload: `map_sizes[m]` target
    points of the large-block generator with scene seed `map_seeds[m]`;
    `source_sizes[m][j]` points for source j of map m."""
    map_seeds: list
    map_sizes: list
    source_sizes: list


def serving_pool(n_maps: int = 32, sources_per_map: int = 8, map_lo: int = 500_000,
                 map_hi: int = 4_000_000, src_lo: int = 2_000, src_hi: int = 200_000) -> ServingPool:
    """32 maps of 0.5M-4M points (log-uniform) from the large-block generator
    (each its own seeded scene) and, per map, request sources log-uniform in
    [2k, 200k] points."""
    rng = Rng(SEED_BASE + 1000 * 3)
    msz = np.exp(rng.uniform(n_maps, np.log(map_lo), np.log(map_hi))).astype(np.int64)
    ssz = np.exp(rng.uniform(n_maps * sources_per_map, np.log(src_lo), np.log(src_hi))).astype(np.int64)
    seeds = [SEED_BASE + 1000 * 3 + 101 * m for m in range(n_maps)]
    return ServingPool(seeds, [int(v) for v in msz],
                       [[int(v) for v in ssz[m * sources_per_map:(m + 1) * sources_per_map]]
                        for m in range(n_maps)])


def serving_map(pool: ServingPool, m: int) -> np.ndarray:
    """Target cloud of map m (sigma 0.02 m)."""
    seed = pool.map_seeds[m]
    return sample_scene(large_block_scene(seed), pool.map_sizes[m], 0.02, seed * 7 + 1)


def serving_source(pool: ServingPool, m: int, j: int) -> Workload:
    """Source j of map m: an independent sample of map m's scene moved by a
    seeded offset of up to 15 deg / 2 m (the large-block generator's); leaf
    0.2, k 15, d_max 2.0, at most 40 iterations."""
    seed = pool.map_seeds[m]
    sseed = seed * 7 + 11 + 13 * j
    src0 = sample_scene(large_block_scene(seed), pool.source_sizes[m][j], 0.02, sseed)
    q, t = random_motion(sseed, 15.0, 2.0, per_axis=False)
    gq, gt = invert_motion(q, t)
    return Workload(f"serve_{m}_{j}", apply_motion(q, t, src0), np.zeros((0, 3), np.float32), gq,
                    gt, 0.2, 15, 2.0, 40, 100.0)
