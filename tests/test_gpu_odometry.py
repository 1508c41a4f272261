"""Scan matching and rolling-map odometry on the HIP device (gpu).

Bars:
  GicpScanMatcher                the reference's IcpScanMatcher test case
                                 (crates/spatialrust-mapping/src/scan.rs:330-356) with GICP:
                                 translation within 5e-3
  rolling map contents           bit-exact vs a CPU restatement of the insertion rule, given
                                 the poses the device returned (voxel filter by the oracle)
  odometry                       deterministic (identical poses on a re-run); drift vs the
                                 ground-truth trajectory below 0.1 m / 0.5 deg over 30 m
  incremental map upkeep         poses bit-identical to recomputing every normal
  full scan size (64 x 2048 beams, ~118k points per scan, 40 scans, 4M-voxel map)
                                 map contents bit-exact vs the insertion rule,
                                 incremental == full recompute, drift below 0.1 m / 0.5 deg
"""
import numpy as np
import pytest

import spatialrust_amd as sr
from spatialrust_amd import synth

pytestmark = pytest.mark.gpu

LEAF = 0.25


def _scans(n, beams=16, bins=512):
    scene = synth.odometry_scene(synth.SEED_BASE + 4000)
    poses = synth.odometry_trajectory(n, seed=1)
    scans = [synth.lidar_scan(scene, q, t, seed=100 + k, beams=beams, azimuth_bins=bins)
             for k, (q, t) in enumerate(poses)]
    return scans, poses


def _cfg():
    return sr.GicpConfig(max_iterations=30, k_neighbors=15, max_correspondence_distance=1.0)


def _rel(pose0, pose):
    """pose0^-1 o pose for (q, t) pairs, FP64."""
    q0, t0 = pose0
    q0i, t0i = synth.invert_motion(q0, t0)
    R = synth.quat_to_mat(q0i)
    q = _qmul(q0i, pose[0])
    return q, R @ np.asarray(pose[1]) + t0i


def _qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def _quat_mat_device_order(q):
    # the FP64 rotation the device builds from its quaternion (same operation order)
    x, y, z, w = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _expected_map(orc, scans, results, leaf, capacity):
    """Insertion rule restated on the CPU: voxel-filter each scan (oracle),
    p' = ((r0 x + r1 y) + r2 z) + t in FP64 rounded to f32, key = floor(p' *
    f32(1/leaf)); new keys appended in point order (lowest index per key);
    while the new voxels do not fit, the oldest map voxels are dropped."""
    inv = np.float32(1.0) / np.float32(leaf)
    lst = np.zeros((0, 3), np.float32)
    for scan, res in zip(scans, results):
        cent, _, _ = orc.voxel_centroid(scan, leaf)
        R = _quat_mat_device_order(res.rotation64)
        t = np.asarray(res.translation64, np.float64)
        p = cent.astype(np.float64)
        w = np.stack([((R[a, 0] * p[:, 0] + R[a, 1] * p[:, 1]) + R[a, 2] * p[:, 2]) + t[a]
                      for a in range(3)], 1).astype(np.float32)
        keys = np.floor(w * inv).astype(np.int64)
        while True:
            have = {tuple(k) for k in np.floor(lst * inv).astype(np.int64)}
            fresh, seen = [], set()
            for i, k in enumerate(map(tuple, keys)):
                if k not in have and k not in seen:
                    seen.add(k)
                    fresh.append(i)
            if len(lst) + len(fresh) <= capacity:
                break
            lst = lst[min(len(lst) + len(fresh) - capacity, len(lst)):]
        lst = np.concatenate([lst, w[fresh]], 0)
    return lst


def test_gicp_scan_matcher_recovers_motion(hip):
    g = np.stack(np.meshgrid(np.arange(6), np.arange(6), np.arange(3), indexing="ij"), -1)
    previous = (g.reshape(-1, 3) * 0.05).astype(np.float32)
    current = previous + np.array([0.02, -0.01, 0.0], np.float32)
    m = sr.GicpScanMatcher(sr.GicpConfig(max_correspondence_distance=0.1, max_iterations=30))
    est = m.match_scans(previous, current)
    assert abs(est.translation[0] - 0.02) < 5e-3
    assert abs(est.translation[1] + 0.01) < 5e-3


def test_odometry_map_is_the_insertion_rule_and_deterministic(hip, orc):
    scans, _ = _scans(12)
    runs = []
    for _ in range(2):
        odo = sr.RollingMapOdometry(_cfg(), voxel_leaf=LEAF, max_map_points=4_000_000)
        res = [odo.push(s) for s in scans]
        runs.append((res, odo.map_cloud()))
        odo.close()
    (res, cloud), (res2, cloud2) = runs
    for a, b in zip(res, res2):
        assert np.array_equal(a.rotation64, b.rotation64) and np.array_equal(a.translation64, b.translation64)
    assert np.array_equal(cloud, cloud2)
    want = _expected_map(orc, scans, res, LEAF, 4_000_000)
    assert cloud.shape == want.shape
    assert np.array_equal(cloud, want)


def test_odometry_rolling_capacity(hip, orc):
    scans, _ = _scans(8)
    cap = 9000
    odo = sr.RollingMapOdometry(_cfg(), voxel_leaf=LEAF, max_map_points=cap)
    res = []
    for s in scans:
        res.append(odo.push(s))
        assert odo.map_points() <= cap
    cloud = odo.map_cloud()
    odo.close()
    want = _expected_map(orc, scans, res, LEAF, cap)
    assert np.array_equal(cloud, want)


def test_odometry_drift(hip):
    scans, poses = _scans(31)
    odo = sr.RollingMapOdometry(_cfg(), voxel_leaf=LEAF)
    res = [odo.push(s) for s in scans]
    odo.close()
    q_gt, t_gt = _rel(poses[0], poses[-1])
    rot = synth.rotation_angle_between(res[-1].rotation64, q_gt)
    trans = float(np.linalg.norm(res[-1].translation64 - t_gt))
    assert trans < 0.1 and rot < np.deg2rad(0.5), (trans, rot)


@pytest.mark.parametrize("capacity", [4_000_000, 9000])
def test_odometry_incremental_map_matches_full_recompute(hip, capacity):
    """The map keeps its grid geometry between full rebuilds and recomputes
    only the normals map_refresh_set names; every other normal then equals a
    full recompute on that geometry, so the poses are bit-identical with the
    incremental upkeep on and off (the capacity case drops old voxels, which
    forces full rebuilds)."""
    scans, _ = _scans(30)
    runs = []
    for inc in (1, 0):
        sr.set_option("odometry_incremental", inc)
        try:
            odo = sr.RollingMapOdometry(_cfg(), voxel_leaf=LEAF, max_map_points=capacity)
            res, counts = [], []
            for s in scans:
                res.append(odo.push(s))
                st = odo.map_stats()
                counts.append((st["refreshed"], st["map_points"]))
            stats = odo.map_stats()
            odo.close()
        finally:
            sr.set_option("odometry_incremental", 1)
        runs.append((res, counts, stats))
    (a, cnt_a, st_a), (b, cnt_b, st_b) = runs
    for k, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x.rotation64, y.rotation64), k
        assert np.array_equal(x.translation64, y.translation64), k
        assert (x.fitness, x.iterations, x.correspondences) == (y.fitness, y.iterations, y.correspondences), k
    assert st_a["rebuilds"] == st_b["rebuilds"] and st_a["scans"] == 30
    assert all(r == n for r, n in cnt_b)
    if capacity == 4_000_000:
        assert st_a["rebuilds"] < 30
        assert any(r < n for r, n in cnt_a), cnt_a


def test_odometry_full_scan_size(hip, orc):
    import concurrent.futures
    import multiprocessing
    n = 40
    # ray casting in a spawned process: torch there is the first to initialise HIP
    with concurrent.futures.ProcessPoolExecutor(
            max_workers=1, mp_context=multiprocessing.get_context("spawn")) as ex:
        scans = ex.submit(synth.odometry_scans, n).result()
    poses = synth.odometry_trajectory(n, seed=1)
    assert min(len(s) for s in scans) > 100_000
    runs = []
    for inc in (1, 0):
        sr.set_option("odometry_incremental", inc)
        try:
            odo = sr.RollingMapOdometry(_cfg(), voxel_leaf=LEAF, max_map_points=4_000_000)
            res = [odo.push(s) for s in scans]
            runs.append((res, odo.map_cloud(), odo.map_stats()))
            odo.close()
        finally:
            sr.set_option("odometry_incremental", 1)
    (res, cloud, st), (res0, cloud0, _) = runs
    for k, (a, b) in enumerate(zip(res, res0)):
        assert np.array_equal(a.rotation64, b.rotation64) and np.array_equal(a.translation64, b.translation64), k
    assert np.array_equal(cloud, cloud0)
    assert st["rebuilds"] < n
    want = _expected_map(orc, scans, res, LEAF, 4_000_000)
    assert np.array_equal(cloud, want)
    q_gt, t_gt = _rel(poses[0], poses[-1])
    rot = synth.rotation_angle_between(res[-1].rotation64, q_gt)
    trans = float(np.linalg.norm(res[-1].translation64 - t_gt))
    assert trans < 0.1 and rot < np.deg2rad(0.5), (trans, rot)


def test_odometry_push_many_equals_push(hip):
    """Several odometry streams pushed together (one batched registration per
    step) give each stream the poses and map that pushing it alone gives."""
    scans, _ = _scans(30)
    streams = [scans[0:10], scans[10:20], scans[20:30]]
    alone = []
    for st in streams:
        odo = sr.RollingMapOdometry(_cfg(), voxel_leaf=LEAF)
        alone.append(([odo.push(s) for s in st], odo.map_cloud()))
        odo.close()
    odos = [sr.RollingMapOdometry(_cfg(), voxel_leaf=LEAF) for _ in streams]
    got = [[] for _ in streams]
    for k in range(10):
        for j, r in enumerate(sr.RollingMapOdometry.push_many(odos, [st[k] for st in streams])):
            got[j].append(r)
    for j, (res, cloud) in enumerate(alone):
        for k, (a, b) in enumerate(zip(res, got[j])):
            assert np.array_equal(a.rotation64, b.rotation64), (j, k)
            assert np.array_equal(a.translation64, b.translation64), (j, k)
            assert (a.fitness, a.iterations, a.source_points, a.target_points) == \
                (b.fitness, b.iterations, b.source_points, b.target_points), (j, k)
        assert np.array_equal(cloud, odos[j].map_cloud()), j
    with pytest.raises(sr.InvalidArgument, match="twice"):
        sr.RollingMapOdometry.push_many([odos[0], odos[0]], [streams[0][0], streams[0][1]])
    for o in odos:
        o.close()
