"""HIP path at the benchmark configurations and on the claims that a single
pass cannot show (gpu).

  large block (2M-point target, 200k source; leaf 0.2, k 15, d_max 2.0, 50
    iterations)           pose vs the FP64 oracle golden: <= 1e-4 rad and
                          <= 1 mm; voxel counts exact; fitness within 1e-4 relative
  256-pair batch (small-cube generator, sources log-uniform in [5k, 120k])
                          batched == one-at-a-time, bit for bit; eight golden
                          instances vs the FP64 oracle at <= 1e-4 rad, <= 1 mm
  certified match reuse   on == off, bit for bit (poses, fitness, iterations;
                          on the batch, on the large block, and on a partial
                          overlap whose misses are kept under budgets too),
                          and on does less search work
  pass loops as HIP graphs (eager, captured, replayed) == eager, bit for bit
  singular normal equations (gicp.rs:190-194, linalg.rs:25-27)
                          the reference's error on an input whose H has an
                          exactly zero row
  radius covariances      pose vs the FP64 oracle fed the oracle's radius
                          covariances: <= 1e-4 rad and <= 1 mm
  two source points per thread (2.3M points)
                          alone == inside a batch under every launch
                          candidate, bit for bit; deterministic
  serving workload (the three smallest of its 32 maps, 0.50M-0.54M points, all
    their 24 sources of 2k-200k points; leaf 0.2, k 15, d_max 2.0, 40
    iterations or a pose update below 1e-6)
                          served while others iterate == alone, bit for bit;
                          the golden requests vs the FP64 oracle at <= 1e-4
                          rad, <= 1 mm
"""
import ctypes
import json
import os

import numpy as np
import pytest

import spatialrust_amd as sr
from spatialrust_amd import devmem, synth

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixed(**kw):
    return sr.GicpConfig(transformation_epsilon=0.0, fitness_epsilon=0.0, **kw)


def _pose_err(res, q, t):
    return (synth.rotation_angle_between(res.rotation64, q),
            float(np.linalg.norm(res.translation64 - np.asarray(t))))


def _same(a, b):
    return (np.array_equal(a.rotation64, b.rotation64) and
            np.array_equal(a.translation64, b.translation64) and a.fitness == b.fitness and
            a.iterations == b.iterations and a.converged == b.converged)


def test_large_block_matches_fp64_golden(hip):
    w = synth.large_block_pair()
    ref = json.load(open(os.path.join(GOLDEN, "fp64_poses.json")))[w.name]
    print(f"large block: {len(w.source)} source, {len(w.target)} target points", flush=True)
    src, tgt = devmem.DeviceCloud(w.source), devmem.DeviceCloud(w.target)
    try:
        cfg = _fixed(max_iterations=w.max_iterations, k_neighbors=w.k_neighbors,
                     max_correspondence_distance=w.max_correspondence_distance,
                     voxel_leaf=w.voxel_leaf)
        r = sr.GicpRegistration(cfg).align(src.view, tgt.view)
        host = sr.GicpRegistration(cfg).align(w.source, w.target)  # host views: same result
    finally:
        src.close()
        tgt.close()
    assert (r.source_points, r.target_points) == (182352, 897684)
    assert (r.source_points, r.target_points) == (ref["source_points"], ref["target_points"])
    assert r.iterations == ref["iterations"] == 50
    rot, tr = _pose_err(r, ref["q"], ref["t"])
    print(f"large block vs FP64 oracle: {rot:.3e} rad, {tr:.3e} m", flush=True)
    assert rot <= 1e-4 and tr <= 1e-3 and tr <= 1e-4 * w.extent
    assert abs(r.fitness - ref["fitness"]) <= 1e-4 * ref["fitness"]
    assert _same(r, host)


@pytest.fixture(scope="module")
def batch256():
    pairs = synth.batch_pairs(count=256)
    srcs = [devmem.DeviceCloud(p.source) for p in pairs]
    tgts = [devmem.DeviceCloud(p.target) for p in pairs]
    yield pairs, srcs, tgts
    for c in srcs + tgts:
        c.close()


BATCH_CFG = dict(max_iterations=30, k_neighbors=20, max_correspondence_distance=1.0, voxel_leaf=0.1)


def test_batch_256_equals_plain_and_fp64_golden(hip, batch256):
    pairs, srcs, tgts = batch256
    cfg = _fixed(**BATCH_CFG)
    gold = json.load(open(os.path.join(GOLDEN, "batch_poses.json")))
    assert gold["count"] == len(pairs) == 256
    batched = sr.batch_align_pairs([c.view for c in srcs], [c.view for c in tgts], cfg)
    reg = sr.GicpRegistration(cfg)
    for i in range(len(pairs)):
        alone = reg.align(srcs[i].view, tgts[i].view)
        assert _same(alone, batched[i]), i
    worst = (0.0, 0.0)
    for key, ref in gold["instances"].items():
        i = int(key)
        r = batched[i]
        assert len(pairs[i].source) == ref["raw_source_points"]
        assert (r.source_points, r.target_points) == (ref["source_points"], ref["target_points"])
        rot, tr = _pose_err(r, ref["q"], ref["t"])
        worst = (max(worst[0], rot), max(worst[1], tr))
        assert rot <= 1e-4 and tr <= 1e-3, (i, rot, tr)
        assert abs(r.fitness - ref["fitness"]) <= 1e-4 * ref["fitness"], i
    print(f"batch golden instances: worst {worst[0]:.3e} rad, {worst[1]:.3e} m", flush=True)


def _search_points(fn):
    """Runs fn() with search statistics on; returns (its result, points tested
    by the Gauss-Newton passes)."""
    from spatialrust_amd import native
    lib = native.load()
    st = (ctypes.c_uint64 * 10)()
    lib.sr_profile_read_stats(b"gicp_pass", st)  # clears
    lib.sr_profile_enable(2)
    try:
        out = fn()
    finally:
        lib.sr_profile_enable(0)
    lib.sr_profile_read_stats(b"gicp_pass", st)
    lib.sr_profile_reset()
    return out, int(st[3])


@pytest.fixture
def reuse_restored():
    yield
    sr.set_option("match_reuse", 1)


def test_match_reuse_off_is_bit_identical(hip, reuse_restored, batch256):
    assert sr.get_option("match_reuse") == 1
    w = synth.small_cube_pair()
    pairs, srcs, tgts = batch256
    # a subset of over 4096 blocks of 256 source points: the batch runs the
    # two-kernel first stage (settle + marked search)
    sub = list(range(0, 256, 3))
    assert sum(len(pairs[i].source) for i in sub) * 0.85 > 4096 * 256
    runs = {}
    for reuse in (1, 0):
        sr.set_option("match_reuse", reuse)
        fixed = sr.GicpRegistration(_fixed(max_iterations=30, voxel_leaf=0.1))
        dflt = sr.GicpRegistration(sr.GicpConfig(max_iterations=30, voxel_leaf=0.1))
        a, pts = _search_points(lambda: fixed.align(w.source, w.target))
        b = dflt.align(w.source, w.target)
        c = sr.batch_align_pairs([srcs[i].view for i in sub], [tgts[i].view for i in sub],
                                 _fixed(**BATCH_CFG))
        runs[reuse] = (a, b, c, pts)
    on, off = runs[1], runs[0]
    assert _same(on[0], off[0])
    assert _same(on[1], off[1])
    assert all(_same(x, y) for x, y in zip(on[2], off[2]))
    print(f"points tested over 31 passes: reuse on {on[3]}, off {off[3]}", flush=True)
    assert on[3] < off[3]


@pytest.fixture
def graphs_restored():
    yield
    sr.set_option("pass_graphs", 1)


def test_pass_graphs_off_is_bit_identical(hip, graphs_restored, batch256):
    """A repeated launch shape runs eagerly, is captured as a HIP graph on its
    second run and replayed after that; with pass_graphs 0 every loop runs
    eagerly. The same registrations (one pair alone, and a 6-pair batch) give
    the same bits every time either way."""
    assert sr.get_option("pass_graphs") == 1
    w = synth.small_cube_pair()
    src, tgt = devmem.DeviceCloud(w.source), devmem.DeviceCloud(w.target)
    pairs, srcs, tgts = batch256
    sub = [3, 17, 40, 88, 150, 201]
    reg = sr.GicpRegistration(_fixed(max_iterations=30, voxel_leaf=0.1))
    runs = {}
    try:
        for graphs in (1, 0):
            sr.set_option("pass_graphs", graphs)
            one = [reg.align(src.view, tgt.view) for _ in range(3)]
            many = [sr.batch_align_pairs([srcs[i].view for i in sub], [tgts[i].view for i in sub],
                                         _fixed(**BATCH_CFG)) for _ in range(3)]
            runs[graphs] = (one, many)
    finally:
        src.close()
        tgt.close()
    ref = runs[1][0][0]
    for graphs in (1, 0):
        one, many = runs[graphs]
        assert all(_same(ref, r) for r in one), graphs
        assert all(_same(a, b) for m in many for a, b in zip(runs[1][1][0], m)), graphs


def test_match_reuse_off_is_bit_identical_partial_overlap(hip, reuse_restored):
    """A target cropped to 60% of the cube in x leaves about a third of the
    source with no target point within d_max on every pass (box and
    empty-space rejects, fruitless ring searches): reuse keeps those misses
    under a movement budget too, and the result must still be the searched
    one, bit for bit."""
    w = synth.small_cube_pair()
    tgt = w.target[w.target[:, 0] < np.quantile(w.target[:, 0], 0.6)]
    runs = {}
    for reuse in (1, 0):
        sr.set_option("match_reuse", reuse)
        reg = sr.GicpRegistration(_fixed(max_iterations=30, voxel_leaf=0.1))
        runs[reuse] = _search_points(lambda: reg.align(w.source, tgt))
    (on, pts_on), (off, pts_off) = runs[1], runs[0]
    assert _same(on, off)
    assert on.correspondences == off.correspondences
    misses = on.source_points - on.correspondences
    print(f"partial overlap: {misses} of {on.source_points} source points without a match; "
          f"points tested over 31 passes: reuse on {pts_on}, off {pts_off}", flush=True)
    assert misses > on.source_points // 5
    assert pts_on < pts_off


def test_match_reuse_off_is_bit_identical_large_block(hip, reuse_restored):
    """Over the large block's 32 converged passes the pose moves points by far
    less than an f32 ulp: reuse keeps a match whose point is back at the very
    coordinates of its search even when its movement budget is nil, and the
    result must still be the searched one, bit for bit."""
    w = synth.large_block_pair()
    src, tgt = devmem.DeviceCloud(w.source), devmem.DeviceCloud(w.target)
    runs = {}
    try:
        cfg = _fixed(max_iterations=w.max_iterations, k_neighbors=w.k_neighbors,
                     max_correspondence_distance=w.max_correspondence_distance,
                     voxel_leaf=w.voxel_leaf)
        for reuse in (1, 0):
            sr.set_option("match_reuse", reuse)
            reg = sr.GicpRegistration(cfg)
            runs[reuse] = _search_points(lambda: reg.align(src.view, tgt.view))
    finally:
        src.close()
        tgt.close()
    (on, pts_on), (off, pts_off) = runs[1], runs[0]
    assert _same(on, off)
    assert on.correspondences == off.correspondences
    print(f"large block: points tested over 51 passes: reuse on {pts_on}, off {pts_off}",
          flush=True)
    assert pts_on < pts_off


def test_singular_normal_equations(hip, orc):
    """Source and target on the x axis: with the identity pose every
    transformed point has y = z = 0 exactly, so the rotation about x moves no
    point, the first row and column of H are exactly zero, and both the
    reference's partial-pivot elimination and the LDL^T stop on that pivot."""
    x = np.arange(200, dtype=np.float32) * np.float32(0.05)
    z0 = np.zeros_like(x)
    tgt = np.stack([x, z0, z0], 1)
    src = np.stack([x + np.float32(0.01), z0, z0], 1)
    cfg = sr.GicpConfig(k_neighbors=10, max_iterations=10)
    for mode in ("faithful", "fp64"):
        with pytest.raises(ValueError, match="GICP normal equations were singular"):
            orc.gicp_align(src, tgt, cfg._c(), mode=mode)
    with pytest.raises(sr.InvalidArgument, match="^GICP normal equations were singular$"):
        sr.GicpRegistration(cfg).align(src, tgt)


def _cov6_to_9(c6):
    m = np.zeros((len(c6), 3, 3))
    m[:, 0, 0], m[:, 1, 1], m[:, 2, 2] = c6[:, 0], c6[:, 1], c6[:, 2]
    m[:, 0, 1] = m[:, 1, 0] = c6[:, 3]
    m[:, 0, 2] = m[:, 2, 0] = c6[:, 4]
    m[:, 1, 2] = m[:, 2, 1] = c6[:, 5]
    return m


def test_radius_covariance_pose_matches_fp64_oracle(hip, orc):
    """covariance_radius set (gicp.rs:91-100 GPU covariance branch): the pose
    vs the FP64 oracle loop fed the oracle's radius covariances (the CPU
    evaluation of COV_GRID_WGSL, bit-exact to the device's)."""
    w = synth.small_cube_pair(n_source=30000, n_target=30000)
    sv, _, _ = orc.voxel_centroid(w.source, 0.1)
    tv, _, _ = orc.voxel_centroid(w.target, 0.1)
    radius = 0.3
    sc, _ = orc.radius_covariances(sv, radius, 1e-3)
    tc, _ = orc.radius_covariances(tv, radius, 1e-3)
    cfg = _fixed(max_iterations=20, covariance_radius=radius)
    ref = orc.gicp_align(sv, tv, cfg._c(), mode="fp64", src_cov=_cov6_to_9(sc),
                         tgt_cov=_cov6_to_9(tc))
    r = sr.GicpRegistration(cfg).align(sv, tv)
    assert r.iterations == ref["iterations"] == 20
    rot, tr = _pose_err(r, ref["q"], ref["t"])
    print(f"radius covariances: {rot:.3e} rad, {tr:.3e} m vs FP64", flush=True)
    assert rot <= 1e-4 and tr <= 1e-3
    assert abs(r.fitness - ref["fitness"]) <= 1e-4 * ref["fitness"]


def test_map_rejects_mismatched_config(hip):
    w = synth.small_cube_pair(n_source=5000, n_target=8000)
    base = sr.GicpConfig(voxel_leaf=0.1, max_iterations=5)
    m = sr.TargetMap(w.target, base)
    try:
        m.align(w.source, sr.GicpConfig(voxel_leaf=0.1, max_iterations=3))  # iteration cap may differ
        for bad in (sr.GicpConfig(voxel_leaf=0.1, epsilon=1e-2), sr.GicpConfig(voxel_leaf=0.2),
                    sr.GicpConfig(voxel_leaf=0.1, k_neighbors=12),
                    sr.GicpConfig(voxel_leaf=0.1, covariance_radius=0.3)):
            with pytest.raises(sr.InvalidArgument, match="must match the map"):
                m.align(w.source, bad)
            with pytest.raises(sr.InvalidArgument, match="must match the map"):
                sr.batch_align([m], [w.source], bad)
    finally:
        m.close()


def test_k_above_network_size_is_rejected_before_work(hip):
    cloud = synth.small_cube_pair(n_source=3000, n_target=3000).target
    with pytest.raises(sr.InvalidArgument, match="k_neighbors <= 64"):
        sr.GicpRegistration(sr.GicpConfig(k_neighbors=65)).align(cloud, cloud)
    with pytest.raises(sr.InvalidArgument, match="k_neighbors <= 64"):
        sr.Server(sr.GicpConfig(k_neighbors=80), hbm_budget_bytes=1 << 30, max_active=2)
    # the radius path does not use k
    r = sr.GicpRegistration(sr.GicpConfig(k_neighbors=65, covariance_radius=0.3,
                                          max_iterations=2)).align(cloud, cloud)
    assert r.iterations >= 1


@pytest.fixture
def tiles_restored():
    yield
    sr.set_option("radius_tiles", 0)


@pytest.mark.parametrize("case", ["large_block", "dense"])
def test_radius_tiles_equal_per_point_form(hip, orc, tiles_restored, case):
    """Radius covariances and grid normals from LDS-staged cell tiles equal the
    per-point form bit for bit: on a 300k-point sample of the large block
    (most tiles staged), and on a dense cloud whose neighbourhoods exceed the
    LDS halo (the tiles fall back to device-memory reads), which also matches
    the oracle's evaluation of COV_GRID_WGSL."""
    rng = np.random.default_rng(21)
    if case == "large_block":
        pts = synth.large_block_pair(n_source=1000, n_target=300_000).target
        radius = 0.4
    else:
        pts = (rng.random((20000, 3)) * [1.0, 1.0, 0.2]).astype(np.float32)
        radius = 0.25
    out = {}
    for tiles in (1, 0):
        sr.set_option("radius_tiles", tiles)
        cov, nrm = sr.plane_covariances(pts, radius=radius)
        est = sr.GpuNormalEstimator(sr.NormalEstimationConfig(search_radius=radius,
                                                              viewpoint=(0.0, 0.0, 30.0)))
        out[tiles] = (cov, nrm) + tuple(est.estimate(pts))
    for a, b in zip(out[1], out[0]):
        assert np.array_equal(a, b)
    if case == "dense":
        o, _ = orc.radius_covariances(pts, radius, 1e-3)
        assert np.array_equal(out[1][0], o)


@pytest.fixture
def tune_restored():
    yield
    sr.set_option("tune_force", -1)


def test_autotune_candidates_bit_identical(hip, tune_restored):
    """Every launch-tuning candidate gives the pose, fitness and iteration
    count of the tuner's own choice, bit for bit: for one registration large
    enough to be tuned (over 128k source points after the voxel filter), and for
    a batch of ~1.8M source points, where the candidates run the first stage
    as two kernels or as one."""
    w = synth.small_cube_pair(n_source=200000, n_target=200000)
    cfg = _fixed(max_iterations=20, voxel_leaf=0.1)
    reg = sr.GicpRegistration(cfg)
    ref = reg.align(w.source, w.target)
    assert ref.source_points >= 1 << 17
    assert "single" in sr.tune_describe()
    for c in range(sr.get_option("tune_candidates_single")):
        sr.set_option("tune_force", c)
        assert _same(reg.align(w.source, w.target), ref), c
    sr.set_option("tune_force", -1)
    pairs = synth.batch_pairs(count=40, lo=30000, hi=60000, seed_offset=13)
    srcs = [devmem.DeviceCloud(p.source) for p in pairs]
    tgts = [devmem.DeviceCloud(p.target) for p in pairs]
    try:
        sv, tv = [c.view for c in srcs], [c.view for c in tgts]
        bref = sr.batch_align_pairs(sv, tv, cfg)
        assert "batch" in sr.tune_describe()
        for c in range(sr.get_option("tune_candidates_batch")):
            sr.set_option("tune_force", c)
            got = sr.batch_align_pairs(sv, tv, cfg)
            assert all(_same(a, b) for a, b in zip(got, bref)), c
    finally:
        for c in srcs + tgts:
            c.close()
    print(sr.tune_describe(), flush=True)


def test_two_points_per_thread_stages_agree(hip, tune_restored):
    """Clouds above 2^21 points run two source points per thread, and the first
    stage then appends to the queue once per point of a thread (the workgroup
    reservation is reused within a launch). A 2.3M-point registration alone
    (split first stage) and twice inside a batch under every batch candidate
    (split and one-kernel first stage) give the same pose, fitness and
    iterations, bit for bit, and the registration alone is deterministic."""
    scene = synth.small_cube_scene(synth.SEED_BASE + 77)
    tgt = synth.sample_scene(scene, 2_300_000, 0.01, 11)
    q, t = synth.random_motion(78, 6.0, 0.3, per_axis=True)
    src = synth.apply_motion(q, t, synth.sample_scene(scene, 2_300_000, 0.01, 12))
    cfg = _fixed(max_iterations=6, k_neighbors=20, max_correspondence_distance=1.0, voxel_leaf=0.0)
    s_dev, t_dev = devmem.DeviceCloud(src), devmem.DeviceCloud(tgt)
    try:
        reg = sr.GicpRegistration(cfg)
        ref = reg.align(s_dev.view, t_dev.view)
        assert ref.source_points >= 1 << 21
        assert _same(reg.align(s_dev.view, t_dev.view), ref)
        for c in range(sr.get_option("tune_candidates_batch")):
            sr.set_option("tune_force", c)
            got = sr.batch_align_pairs([s_dev.view, s_dev.view], [t_dev.view, t_dev.view], cfg)
            assert _same(got[0], ref) and _same(got[1], ref), c
    finally:
        s_dev.close()
        t_dev.close()


def test_serving_workload_maps_served_equal_alone_and_fp64_golden(hip):
    pool = synth.serving_pool()
    ms = [int(m) for m in np.argsort(pool.map_sizes, kind="stable")[:3]]
    cfg = sr.GicpConfig(voxel_leaf=0.2, k_neighbors=15, max_correspondence_distance=2.0,
                        max_iterations=40, transformation_epsilon=1e-6, fitness_epsilon=0.0)
    maps = {m: sr.TargetMap(synth.serving_map(pool, m), cfg) for m in ms}
    reqs = [(m, j) for m in ms for j in range(len(pool.source_sizes[m]))]
    srcs = {mj: synth.serving_source(pool, *mj) for mj in reqs}
    order = np.random.default_rng(7).permutation(len(reqs))
    server = sr.Server(cfg, hbm_budget_bytes=8 << 30, max_active=256, max_iterations=40,
                       delta_pose_epsilon=1e-6)
    tickets = {}
    for k, r in enumerate(order):  # arrivals spread over the steps
        mj = reqs[r]
        tickets[server.submit(maps[mj[0]], srcs[mj].source)] = mj
        if k % 3 == 2:
            server.step()
    steps = 0
    while server.step():
        steps += 1
        assert steps < 2000
    assert sorted(server.finished()) == sorted(tickets)
    gold = json.load(open(os.path.join(GOLDEN, "serving_poses.json")))["requests"]
    checked = 0
    for t, mj in tickets.items():
        got = server.poll(t)
        alone = maps[mj[0]].align(srcs[mj].source, cfg)
        assert _same(got, alone), mj
        g = gold.get(f"{mj[0]}_{mj[1]}")
        if g is not None:
            rot, trans = _pose_err(got, g["q"], g["t"])
            assert rot <= 1e-4 and trans <= 1e-3, (mj, rot, trans)
            assert abs(got.iterations - g["iterations"]) <= 1, (mj, got.iterations)
            checked += 1
    assert checked == len(gold)
    server.close()
    for m in maps.values():
        m.close()
