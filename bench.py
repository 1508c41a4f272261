"""GICP registrations/s on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload large_block|small_cube]

A step is one full registration with both clouds already resident in HBM:
voxel filter of source and target, uniform grids, k-NN plane covariances and a
fixed number of Gauss-Newton passes (transformation_epsilon = fitness_epsilon
= 0, so every iteration runs), ending with the pose read back to the host.
The timed region runs with kernel timing off (the pass loop replays as a HIP
graph); a second run of the same steps with eager, event-bracketed launches
gives the per-kernel times and the roofline figure.

Default workload (single large instance): 2M-point target, 200k-point source,
100 x 100 x 10 m block, voxel leaf 0.2 m, k = 15, d_max = 2.0 m, 50 iterations.

One JSON line on stdout; `roofline` is the dominant kernel work, the
Gauss-Newton pass (its stage kernels k_gicp_search_fused, k_gicp_hard and
k_gicp_terms, bracketed together; launches of 4096 or more blocks split the
first into k_gicp_settle and k_gicp_search), timed with HIP events on the stream it runs
on; `cpu_baseline` is the
CPU oracle (f32-faithful restatement of the reference, one thread) on a bounded
sample of the same workload.

The batched leg (`batched_vs_plain`: 256 small-cube pairs, one call against one
registration per call) carries its own `roofline` (72 B per source point per
pass over the library profiler's pass time) and `call_roofline` (the
compulsory bytes of every stage over the timed call). `--no-graphs` keeps
every pass loop on eager launches, for runs under an external profiler.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "GICP registrations/s + p50/p99 latency at ≤1e-4 rad error, batched vs plain"


def _workload(name):
    from spatialrust_amd import synth
    if name == "large_block":
        return synth.large_block_pair()
    if name == "small_cube":
        return synth.small_cube_pair()
    raise SystemExit(f"unknown workload {name}")


def _config(w):
    import spatialrust_amd as sr
    return sr.GicpConfig(max_iterations=w.max_iterations, k_neighbors=w.k_neighbors,
                         max_correspondence_distance=w.max_correspondence_distance,
                         transformation_epsilon=0.0, fitness_epsilon=0.0, voxel_leaf=w.voxel_leaf)


def _cpu_baseline(w, all_core_threads):
    """Faithful CPU oracle (the f32/f64 restatement of the reference's CPU
    path), timed in full on this host: one registration of the workload (voxel
    filter of both clouds, KD-trees, k-NN covariances and every Gauss-Newton
    iteration) on one thread, as the reference runs; the same registration with
    its per-point loops on all cores (informational); and the reference's own
    bench case, the 7,500-point box corner (benches/registration.rs:23-63,
    81-86; published ~26 ms and a 0.00057 m probe error in
    notes/2026-06-15_registration_bench.md:11,36-38)."""
    import oracle
    from spatialrust_amd.registration import GicpConfig
    cfg = GicpConfig(max_iterations=w.max_iterations, k_neighbors=w.k_neighbors,
                     max_correspondence_distance=w.max_correspondence_distance,
                     transformation_epsilon=0.0, fitness_epsilon=0.0)

    def one(threads):
        t0 = time.perf_counter()
        sv, _, _ = oracle.voxel_centroid(w.source, w.voxel_leaf)
        tv, _, _ = oracle.voxel_centroid(w.target, w.voxel_leaf)
        r = oracle.gicp_align(sv, tv, cfg._c(), mode="faithful", threads=threads)
        return time.perf_counter() - t0, r

    t1, r1 = one(1)
    tn, _ = one(all_core_threads)
    corner = _reference_corner_cpu()
    return {"value": 1.0 / t1, "unit": "registrations/s", "cores": 1, "kind": "port",
            "sample": (f"{w.name}: one full registration (voxel filter of both clouds, trees, "
                       f"covariances, {r1['iterations']} Gauss-Newton iterations), measured, "
                       "not extrapolated"),
            "seconds_per_registration": t1,
            "all_core": {"threads": all_core_threads, "seconds_per_registration": tn,
                         "registrations_per_s": 1.0 / tn,
                         "note": "informational: the reference's GICP is single-threaded"},
            "reference_corner": corner,
            "host_cpu": _cpu_model(), "host_nproc": os.cpu_count()}


def _reference_corner_cpu(reps=9):
    """The reference's registration bench case on the CPU oracle: median time
    of `reps` runs and the probe error."""
    import oracle
    from spatialrust_amd.registration import GicpConfig
    from tests import refinputs
    tgt, src, mq, mt = refinputs.bench_corner_pair()
    cfg = GicpConfig(max_correspondence_distance=0.3, max_iterations=40, k_neighbors=12)._c()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r = oracle.gicp_align(src, tgt, cfg, mode="faithful")
        ts.append(time.perf_counter() - t0)
    return {"points": int(len(tgt)), "median_ms": float(np.median(ts) * 1e3),
            "probe_error_m": refinputs.corner_probe_error(r["q"], r["t"], mq, mt),
            "iterations": r["iterations"],
            "published": {"median_ms": 26.0, "probe_error_m": 0.00057,
                          "source": "notes/2026-06-15_registration_bench.md:11,36-38"}}


def _reference_corner_gpu(reps=20):
    """The same corner case through the HIP path (host views, as the
    reference's bench passes its clouds): median time and probe error."""
    import spatialrust_amd as sr
    from tests import refinputs
    tgt, src, mq, mt = refinputs.bench_corner_pair()
    reg = sr.GicpRegistration(sr.GicpConfig(max_correspondence_distance=0.3, max_iterations=40,
                                            k_neighbors=12))
    for _ in range(3):
        reg.align(src, tgt)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r = reg.align(src, tgt)
        ts.append(time.perf_counter() - t0)
    return {"points": int(len(tgt)), "median_ms": float(np.median(ts) * 1e3),
            "iterations": r.iterations, "converged": r.converged,
            "probe_error_m": refinputs.corner_probe_error(r.rotation64, r.translation64, mq, mt)}


def _batched_vs_plain(count):
    """256 independent pairs of the small-cube generator (source sizes
    log-uniform in [5k, 120k], target 1.5x, offsets up to 10 deg / 1 m; leaf 0.1,
    k 20, d_max 1.0, 30 fixed iterations), all clouds resident in HBM. Plain:
    one sr_gicp_align per pair, one after another. Batched: one
    sr_batch_align_pairs call for all pairs. Both legs do the same work (both
    clouds preprocessed, 31 passes) and must return bit-identical poses."""
    import spatialrust_amd as sr
    from spatialrust_amd import devmem, synth
    pairs = synth.batch_pairs(count=count)
    srcs = [devmem.DeviceCloud(p.source) for p in pairs]
    tgts = [devmem.DeviceCloud(p.target) for p in pairs]
    cfg = sr.GicpConfig(max_iterations=30, k_neighbors=20, max_correspondence_distance=1.0,
                        transformation_epsilon=0.0, fitness_epsilon=0.0, voxel_leaf=0.1)
    reg = sr.GicpRegistration(cfg)
    sv, tv = [s.view for s in srcs], [t.view for t in tgts]
    # warm-up of both legs (allocator, graph cache, worker streams); the
    # batched leg's allocator is warmed at its full size by a batch of other
    # pairs of the same generator (so the timed call reuses no graph of its own)
    for s_, t_ in zip(sv[:8], tv[:8]):
        reg.align(s_, t_)
    warm = synth.batch_pairs(count=count, seed_offset=1)
    ws = [devmem.DeviceCloud(p.source) for p in warm]
    wt = [devmem.DeviceCloud(p.target) for p in warm]
    sr.batch_align_pairs([c.view for c in ws], [c.view for c in wt], cfg)
    devmem.synchronize()
    for c in ws + wt:
        c.close()
    del warm
    t0 = time.perf_counter()
    plain = [reg.align(s_, t_) for s_, t_ in zip(sv, tv)]
    devmem.synchronize()
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    batched = sr.batch_align_pairs(sv, tv, cfg)
    devmem.synchronize()
    t_batch = time.perf_counter() - t0
    identical = all(np.array_equal(a.rotation64, b.rotation64) and
                    np.array_equal(a.translation64, b.translation64) for a, b in zip(plain, batched))
    # The same batch once more with the library's profiler on (eager launches,
    # HIP events around each stage on the library's stream): per-pass and
    # preparation stage times for the batched roofline.
    import ctypes
    from spatialrust_amd import native
    lib = native.load()
    lib.sr_profile_reset()
    lib.sr_profile_enable(1)
    again = sr.batch_align_pairs(sv, tv, cfg)
    devmem.synchronize()
    lib.sr_profile_enable(0)
    stage = {}
    for name in ("gicp_pass", "gicp_finalize", "knn_normals", "voxel_insert"):
        ms, cnt = ctypes.c_double(0.0), ctypes.c_uint64(0)
        lib.sr_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        stage[name] = (ms.value, cnt.value)
    lib.sr_profile_reset()
    identical = identical and all(np.array_equal(a.rotation64, b.rotation64) for a, b in zip(batched, again))
    n_src = sum(b.source_points for b in batched)
    n_tgt = sum(b.target_points for b in batched)
    n_raw = sum(len(p.source) + len(p.target) for p in pairs)
    pass_ms, pass_n = stage["gicp_pass"]
    avg_pass_s = pass_ms / max(pass_n, 1) * 1e-3
    pass_bytes = 72 * n_src
    pass_gbs = pass_bytes / avg_pass_s / 1e9 if avg_pass_s > 0 else 0.0
    # whole call, the compulsory bytes of every stage: voxel 12 N + 28 C (both
    # clouds), target grid 20 N_t, covariances 36 N (both clouds), passes 72 N_s
    call_bytes = (12 * n_raw + 28 * (n_src + n_tgt) + 20 * n_tgt + 36 * (n_src + n_tgt)
                  + pass_n * pass_bytes)
    call_gbs = call_bytes / t_batch / 1e9
    for c in srcs + tgts:
        c.close()
    fp64 = None
    gpath = os.path.join(ROOT, "tests", "golden", "batch_poses.json")
    if os.path.exists(gpath):
        gold = json.load(open(gpath))
        if gold["count"] == count:
            rot = [synth.rotation_angle_between(batched[int(i)].rotation64, g["q"])
                   for i, g in gold["instances"].items()]
            tr = [float(np.linalg.norm(batched[int(i)].translation64 - np.array(g["t"])))
                  for i, g in gold["instances"].items()]
            fp64 = {"instances": sorted(int(i) for i in gold["instances"]),
                    "max_rot_rad": float(max(rot)), "max_trans_m": float(max(tr)),
                    "vs": "FP64 CPU oracle (tests/golden/batch_poses.json)"}
    return {"workload": f"batch_{count}_small_cube_pairs", "pairs": count,
            "error_vs_fp64_oracle": fp64,
            "source_points": int(sum(len(p.source) for p in pairs)),
            "target_points": int(sum(len(p.target) for p in pairs)),
            "plain_registrations_per_s": count / t_plain,
            "batched_registrations_per_s": count / t_batch,
            "batched_over_plain": t_plain / t_batch,
            "identical_poses": bool(identical),
            "after_voxel": {"source": int(n_src), "target": int(n_tgt)},
            "stage_ms_per_call": {k: v[0] for k, v in stage.items()},
            "roofline": {"kernel": "gicp pass over all pairs (k_gicp_settle + k_gicp_search + "
                                   "k_gicp_hard + k_gicp_terms)",
                         "bound": "hbm", "achieved": pass_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": pass_gbs / HBM_PEAK_GBS, "traffic": _traffic("gicp_pass", f"batch_{count}_small_cube_pairs"),
                         "algorithmic_bytes_per_launch": pass_bytes,
                         "avg_launch_us": avg_pass_s * 1e6, "launches": pass_n},
            "call_roofline": {"bytes": call_bytes, "achieved": call_gbs, "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": call_gbs / HBM_PEAK_GBS,
                              "formula": "12 N_raw + 28 C + 20 N_t + 36 (N_s + N_t) + passes x 72 N_s"},
            "max_rot_err_vs_ground_truth_rad": float(max(
                synth.rotation_angle_between(b.rotation64, p.gt_q) for b, p in zip(batched, pairs)))}


def _gen_map(m):
    from spatialrust_amd import synth
    return synth.serving_map(synth.serving_pool(), m)


def _gen_source(mj):
    from spatialrust_amd import synth
    w = synth.serving_source(synth.serving_pool(), *mj)
    return w.source, w.gt_q, w.gt_t


def _serving(rates, seconds, max_requests, max_active):
    """Continuous-batching server under Poisson arrivals. 32 device-resident
    maps of 0.5M-4M points (large-block generator, one seeded scene each),
    request sources log-uniform in [2k, 200k] points sampled from the map's
    scene and offset by up to 15 deg / 2 m; leaf 0.2, k 15, d_max 2.0, run until
    the pose update is below 1e-6 or 40 iterations, 200 GB HBM budget. Latency
    is from each request's scheduled arrival to the step that saw it finished;
    throughput is requests over the time from the first arrival to the last
    completion. Each served result is checked bit for bit against the same
    registration run alone (sr_map_align) on a sample of requests, and against
    FP64 oracle poses where tests/golden/serving_poses.json holds one."""
    import concurrent.futures
    import multiprocessing
    import spatialrust_amd as sr
    from spatialrust_amd import synth
    pool = synth.serving_pool()
    nm, spm = len(pool.map_sizes), len(pool.source_sizes[0])
    t0 = time.perf_counter()
    ctx = multiprocessing.get_context("spawn")
    with concurrent.futures.ProcessPoolExecutor(max_workers=min(12, os.cpu_count() or 1),
                                                mp_context=ctx) as ex:
        maps_host = list(ex.map(_gen_map, range(nm)))
        srcs = list(ex.map(_gen_source, [(m, j) for m in range(nm) for j in range(spm)]))
    t_gen = time.perf_counter() - t0
    cfg = sr.GicpConfig(voxel_leaf=0.2, k_neighbors=15, max_correspondence_distance=2.0,
                        max_iterations=40, transformation_epsilon=1e-6, fitness_epsilon=0.0)
    t0 = time.perf_counter()
    maps = [sr.TargetMap(mh, cfg) for mh in maps_host]
    t_maps = time.perf_counter() - t0
    map_points = [m.points() for m in maps]
    del maps_host
    gold = {}
    gpath = os.path.join(ROOT, "tests", "golden", "serving_poses.json")
    if os.path.exists(gpath):
        gold = json.load(open(gpath))["requests"]
    out = {"workload": "serving_32_maps", "maps": nm, "map_points_raw": [int(v) for v in pool.map_sizes],
           "map_points_after_voxel": map_points, "sources_per_map": spm,
           "source_points_raw_mean": float(np.mean([len(s[0]) for s in srcs])),
           "max_active": max_active, "hbm_budget_bytes": 200 << 30, "setup_s": {
               "generate": t_gen, "maps": t_maps}, "rates": []}
    rng = np.random.default_rng(12345)
    for rate, secs in zip(rates, seconds):
        n = int(min(max_requests, rate * secs))
        gaps = rng.exponential(1.0 / rate, n)
        arrive = np.cumsum(gaps) - gaps[0]
        req = [(int(rng.integers(nm)), int(rng.integers(spm))) for _ in range(n)]
        server = sr.Server(cfg, hbm_budget_bytes=200 << 30, max_active=max_active,
                           max_iterations=40, delta_pose_epsilon=1e-6)
        # warm-up: a few requests through the same server
        for k in range(4):
            m, j = req[k]
            server.submit(maps[m], srcs[m * spm + j][0])
        while server.step():
            pass
        for t in server.finished():
            server.poll(t)
        ticket_of = {}
        done_at = np.zeros(n)
        results = {}
        nxt = 0
        ndone = 0
        steps = 0
        peak_bytes = 0
        start = time.perf_counter()
        while ndone < n:
            now = time.perf_counter() - start
            while nxt < n and arrive[nxt] <= now:
                m, j = req[nxt]
                ticket_of[server.submit(maps[m], srcs[m * spm + j][0])] = nxt
                nxt += 1
            pending = server.step()
            steps += 1
            peak_bytes = max(peak_bytes, server.device_bytes())
            fin = server.finished()
            if fin:
                tnow = time.perf_counter() - start
                for t in fin:
                    i = ticket_of.pop(t)
                    results[i] = server.poll(t)
                    done_at[i] = tnow
                    ndone += 1
            if not pending and nxt < n:  # idle until the next arrival
                wait = arrive[nxt] - (time.perf_counter() - start)
                if wait > 0:
                    time.sleep(min(wait, 0.001))
        server.close()
        lat = (done_at - arrive) * 1e3
        span = float(done_at.max() - arrive[0])
        its = [results[i].iterations for i in range(n)]
        gt_rot = max(synth.rotation_angle_between(results[i].rotation64, srcs[req[i][0] * spm + req[i][1]][1])
                     for i in range(n))
        # bit-identity with the same registration alone, on a sample of requests
        sample = list(range(0, n, max(1, n // 6)))[:6]
        ident = True
        for i in sample:
            m, j = req[i]
            alone = maps[m].align(srcs[m * spm + j][0], cfg)
            ident = ident and (np.array_equal(alone.rotation64, results[i].rotation64) and
                               np.array_equal(alone.translation64, results[i].translation64) and
                               alone.iterations == results[i].iterations)
        fp64 = []
        for key, g in gold.items():
            m, j = (int(v) for v in key.split("_"))
            hits = [i for i in range(n) if req[i] == (m, j)]
            if hits:
                r = results[hits[0]]
                fp64.append((synth.rotation_angle_between(r.rotation64, g["q"]),
                             float(np.linalg.norm(r.translation64 - np.array(g["t"])))))
        out["rates"].append({
            "offered_per_s": rate, "requests": n, "achieved_per_s": n / span,
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "max": float(lat.max())},
            "mean_iterations": float(np.mean(its)),
            "converged": float(np.mean([results[i].converged for i in range(n)])),
            "steps": steps, "peak_charged_bytes": int(peak_bytes),
            "max_rot_err_vs_ground_truth_rad": float(gt_rot),
            "alone_equals_served": {"checked": len(sample), "identical": bool(ident)},
            "error_vs_fp64_oracle": ({"checked": len(fp64), "max_rot_rad": max(e[0] for e in fp64),
                                      "max_trans_m": max(e[1] for e in fp64)} if fp64 else None)})
        print(f"serving {rate}/s: {out['rates'][-1]['achieved_per_s']:.0f} reg/s, "
              f"p50 {out['rates'][-1]['latency_ms']['p50']:.1f} ms, "
              f"p99 {out['rates'][-1]['latency_ms']['p99']:.1f} ms", file=sys.stderr, flush=True)
    for m in maps:
        m.close()
    return out


def _radius_leg(w, reps=5):
    """Radius covariances (covariance_radius, COV_GRID_WGSL semantics) and grid
    normals of the workload's target after the voxel filter, radius 0.4 m,
    with the neighbourhoods read from LDS-staged cell tiles (the tile list's
    construction included) and from device memory per point: kernel times (HIP
    events) and bit identity."""
    import ctypes
    import spatialrust_amd as sr
    from spatialrust_amd import native
    lib = native.load()
    vox = sr.VoxelGridDownsample(sr.VoxelGridDownsampleConfig(w.voxel_leaf)).filter_with_policy(
        w.target, sr.ExecutionPolicy.gpu()).points
    radius = 0.4
    est = sr.GpuNormalEstimator(sr.NormalEstimationConfig(search_radius=radius))
    out, res = {}, {}
    for tiles in (1, 0):
        sr.set_option("radius_tiles", tiles)
        cov, _ = sr.plane_covariances(vox, radius=radius)  # warm-up
        lib.sr_profile_reset()
        lib.sr_profile_enable(1)
        for _ in range(reps):
            cov, _ = sr.plane_covariances(vox, radius=radius)
            nrm, curv = est.estimate(vox)
        lib.sr_profile_enable(0)
        t = {}
        for name in ("radius_cov", "grid_normals"):
            ms, cnt = ctypes.c_double(0.0), ctypes.c_uint64(0)
            lib.sr_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
            t[name] = ms.value / max(cnt.value, 1)
        lib.sr_profile_reset()
        out[tiles] = t
        res[tiles] = (cov, nrm, curv)
    sr.set_option("radius_tiles", 0)
    n = len(vox)
    return {"points": int(n), "radius": radius,
            "radius_cov_ms": {"tiles": out[1]["radius_cov"], "per_point": out[0]["radius_cov"]},
            "grid_normals_ms": {"tiles": out[1]["grid_normals"], "per_point": out[0]["grid_normals"]},
            "radius_cov_gbs_per_point": 36.0 * n / (out[0]["radius_cov"] * 1e-3) / 1e9,
            "identical": bool(all(np.array_equal(a, b) for a, b in zip(res[1], res[0])))}


ODO_SCANS = 2000
ODO_STREAMS = 8


def _gen_scans(n):
    """All scans of the odometry leg, ray-cast in FP64 torch on the GPU when
    torch sees one (this runs in a child process), else on the CPU."""
    from spatialrust_amd import synth
    return synth.odometry_scans(n, ODO_SCANS, progress=True)


def _odometry_streams(sr, cfg, scans, streams):
    """The scans cut into `streams` consecutive segments, each its own
    odometry (its own map): pushed one stream after another with
    sr_odometry_push, then all streams in lockstep with sr_odometry_push_many
    (one batched registration per step). Rates are scans over wall time."""
    seg = len(scans) // streams
    parts = [scans[j * seg:(j + 1) * seg] for j in range(streams)]
    alone = []
    t0 = time.perf_counter()
    for part in parts:
        odo = sr.RollingMapOdometry(cfg, voxel_leaf=0.25, max_map_points=4_000_000)
        alone.append([odo.push(s) for s in part])
        odo.close()
    t_alone = time.perf_counter() - t0
    print(f"  {streams} streams one after another: {t_alone:.2f} s", file=sys.stderr, flush=True)
    odos = [sr.RollingMapOdometry(cfg, voxel_leaf=0.25, max_map_points=4_000_000) for _ in parts]
    got = [[] for _ in parts]
    step_ms = []
    t0 = time.perf_counter()
    for k in range(seg):
        t1 = time.perf_counter()
        for j, r in enumerate(sr.RollingMapOdometry.push_many(odos, [p[k] for p in parts])):
            got[j].append(r)
        step_ms.append((time.perf_counter() - t1) * 1e3)
    t_many = time.perf_counter() - t0
    for o in odos:
        o.close()
    same = all(np.array_equal(a.rotation64, b.rotation64) and np.array_equal(a.translation64, b.translation64)
               for ra, rb in zip(alone, got) for a, b in zip(ra, rb))
    step = np.array(step_ms[1:])
    return {"streams": streams, "scans_per_stream": seg,
            "sequential_scans_per_s": streams * seg / t_alone,
            "batched_scans_per_s": streams * seg / t_many,
            "batched_over_sequential": t_alone / t_many,
            "batched_step_ms": {"p50": float(np.percentile(step, 50)), "p99": float(np.percentile(step, 99))},
            "identical_poses": bool(same)}


def _odometry(n_scans):
    """Scan-to-map odometry at full scan size: a spinning LiDAR of 64 beams
    (elevations -25..+15 deg) x 2048 azimuth bins ray-cast into the large-block
    scene (range noise 0.02 m, 5% dropout, ~120k points per scan) along a
    smooth 10 m/s loop at 10 Hz; each scan voxel-filtered at 0.25 m and
    registered (k 15, d_max 1.0, 30 iterations) against the rolling
    device-resident voxel map (4M voxels at most), which it then extends.
    Per-scan latency is the wall time of sr_odometry_push (host scan in, pose
    out); drift is the estimated pose relative to the first scan against the
    ground truth. The same scans then run as ODO_STREAMS streams, batched."""
    import concurrent.futures
    import ctypes
    import multiprocessing
    import spatialrust_amd as sr
    from spatialrust_amd import native, synth
    lib = native.load()
    t0 = time.perf_counter()
    ctx = multiprocessing.get_context("spawn")
    with concurrent.futures.ProcessPoolExecutor(max_workers=1, mp_context=ctx) as ex:
        scans = ex.submit(_gen_scans, n_scans).result()
    t_gen = time.perf_counter() - t0
    poses = synth.odometry_trajectory(ODO_SCANS, seed=1)[:n_scans]
    cfg = sr.GicpConfig(max_iterations=30, k_neighbors=15, max_correspondence_distance=1.0)
    odo = sr.RollingMapOdometry(cfg, voxel_leaf=0.25, max_map_points=4_000_000)
    lat, res, refreshed = [], [], []
    lib.sr_profile_reset()
    lib.sr_profile_enable(1)
    for k in range(n_scans):
        t1 = time.perf_counter()
        res.append(odo.push(scans[k]))
        lat.append(time.perf_counter() - t1)
        st = odo.map_stats()
        refreshed.append(st["refreshed"] / max(st["map_points"], 1))
        if (k + 1) % 250 == 0:
            print(f"  scans registered: {k + 1}/{n_scans}, map {odo.map_points()} voxels",
                  file=sys.stderr, flush=True)
    lib.sr_profile_enable(0)
    spans = {}
    for name in ("odometry_register", "odometry_map"):
        ms, cnt = ctypes.c_double(0.0), ctypes.c_uint64(0)
        lib.sr_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        spans[name] = ms.value / max(cnt.value, 1)
    lib.sr_profile_reset()
    map_points = odo.map_points()
    rebuilds = odo.map_stats()["rebuilds"]
    odo.close()
    q0i, t0i = synth.invert_motion(*poses[0])
    R0i = synth.quat_to_mat(q0i)
    terr, rerr = [], []
    for k in range(n_scans):
        gq = synth.quat_mul(q0i, poses[k][0])
        gt = R0i @ np.asarray(poses[k][1]) + t0i
        terr.append(float(np.linalg.norm(res[k].translation64 - gt)))
        rerr.append(synth.rotation_angle_between(res[k].rotation64, gq))
    path = float(sum(np.linalg.norm(np.asarray(poses[k + 1][1]) - np.asarray(poses[k][1]))
                     for k in range(n_scans - 1)))
    lat_ms = np.array(lat[1:]) * 1e3  # the first scan only seeds the map
    return {"workload": "lidar_64x2048_odometry", "scans": n_scans,
            "points_per_scan_mean": float(np.mean([len(s) for s in scans])),
            "scan_voxels_mean": float(np.mean([r.source_points for r in res[1:]])),
            "map_points_final": int(map_points), "map_capacity": 4_000_000,
            "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99)),
                           "max": float(lat_ms.max())},
            "scans_per_s": float(1e3 / np.mean(lat_ms)),
            "gpu_span_ms_per_scan": {"register": spans["odometry_register"], "map_update": spans["odometry_map"]},
            "map_upkeep": {"full_rebuilds": int(rebuilds),
                           "normals_recomputed_frac_mean": float(np.mean(refreshed[1:])),
                           "normals_recomputed_frac_p50": float(np.median(refreshed[1:]))},
            "drift": {"path_m": path, "final_trans_m": terr[-1], "max_trans_m": max(terr),
                      "final_rot_rad": rerr[-1], "max_rot_rad": max(rerr)},
            "batched_streams": _odometry_streams(sr, cfg, scans, ODO_STREAMS) if n_scans >= 2 * ODO_STREAMS else None,
            "generate_s": t_gen}


def _host_view_rate(w, cfg, steps):
    """The drop-in call as a Rust caller makes it, with host (numpy) clouds:
    the raw xyz columns (12 B per point) are copied to the device inside every
    registration. Not the headline (which has the clouds resident in HBM)."""
    import spatialrust_amd as sr
    reg = sr.GicpRegistration(cfg)
    reg.align(w.source, w.target)
    t0 = time.perf_counter()
    for _ in range(steps):
        reg.align(w.source, w.target)
    dt = (time.perf_counter() - t0) / steps
    return {"registrations_per_s": 1.0 / dt, "ms_per_registration": dt * 1e3, "steps": steps,
            "h2d_bytes_per_registration": 12 * (len(w.source) + len(w.target))}


def _traffic(kernel, workload):
    """HBM bytes per launch of `kernel` (per pass for the pass stage kernels) on
    `workload` from the committed counter summary (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), or None when no summary exists."""
    path = os.path.join(ROOT, "profiles", f"{kernel}.{workload}.pmc.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="large_block")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--search-stats", action="store_true",
                    help="count grid-search work (rings/cells/points per query) on one extra run")
    ap.add_argument("--pass-times", action="store_true",
                    help="print each Gauss-Newton pass duration of the last registration")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-vs-plain leg (256 independent small-cube pairs)")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--no-extra-legs", action="store_true",
                    help="skip the host-view rate and the reference-corner GPU timing")
    ap.add_argument("--no-odometry", action="store_true",
                    help="skip the LiDAR odometry leg (64 x 2048-beam scans, rolling 4M map)")
    ap.add_argument("--odometry-scans", type=int, default=ODO_SCANS)
    ap.add_argument("--no-radius", action="store_true",
                    help="skip the radius covariance / grid normal leg (LDS tiles vs per point)")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving leg (Poisson arrivals against 32 resident maps)")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[200.0, 1000.0, 4000.0])
    ap.add_argument("--serve-seconds", type=float, nargs="+", default=[5.0, 3.0, 1.0])
    ap.add_argument("--serve-max-requests", type=int, default=4000)
    ap.add_argument("--serve-max-active", type=int, default=256)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (a plain copy "
                         "kernel of known size for calibrating hardware byte counters)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="keep every pass loop on eager launches (pass_graphs option 0), for "
                         "runs under an external profiler")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (independent "
              "registrations shard across GPUs with no communication, so one GPU's rate is "
              "the figure)", file=sys.stderr)
        sys.exit(2)

    import spatialrust_amd as sr
    from spatialrust_amd import devmem, native, synth

    if not sr.device_available():
        print("no gfx950 HIP device: " + native.last_error(), file=sys.stderr)
        sys.exit(1)
    lib = native.load()
    if args.no_graphs:
        sr.set_option("pass_graphs", 0)
    if args.copy_mib > 0:
        devmem.copy_calibration(args.copy_mib << 20)
    w = _workload(args.workload)
    cfg = _config(w)
    src = devmem.DeviceCloud(w.source)
    tgt = devmem.DeviceCloud(w.target)
    devmem.synchronize()
    reg = sr.GicpRegistration(cfg)
    for _ in range(args.warmup):
        res = reg.align(src.view, tgt.view)
        print(f"warmup: {res}", file=sys.stderr, flush=True)

    # Timed region: profiling off, so each registration's Gauss-Newton loop
    # replays as one captured HIP graph.
    lat = []
    devmem.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        res = reg.align(src.view, tgt.view)
        lat.append(time.perf_counter() - t0)
    devmem.synchronize()
    total = time.perf_counter() - t_start

    # Kernel-timing region: the same steps again with eager launches, each
    # bracketed by HIP events on the library's stream.
    lib.sr_profile_reset()
    lib.sr_profile_enable(1)
    t_k = time.perf_counter()
    for _ in range(args.steps):
        reg.align(src.view, tgt.view)
    devmem.synchronize()
    total_eager = time.perf_counter() - t_k
    lib.sr_profile_enable(0)

    import ctypes
    kern = {}
    for name in ("gicp_pass", "gicp_finalize", "knn_normals", "voxel_insert"):
        ms = ctypes.c_double(0.0)
        cnt = ctypes.c_uint64(0)
        lib.sr_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        kern[name] = (ms.value, cnt.value)
    pass_ms, pass_n = kern["gicp_pass"]
    if args.pass_times and pass_n:
        each = (ctypes.c_double * pass_n)()
        lib.sr_profile_read_each(b"gicp_pass", each, pass_n)
        per = w.max_iterations + 1
        last = [round(each[i] * 1e3, 1) for i in range(pass_n - per, pass_n)]
        print(f"per-pass us (last registration): {last}", file=sys.stderr, flush=True)
    avg_pass_s = (pass_ms / max(pass_n, 1)) * 1e-3
    algo_bytes = 72 * res.source_points  # source xyz + cov, gathered target xyz + cov
    achieved = algo_bytes / avg_pass_s / 1e9 if avg_pass_s > 0 else 0.0

    if args.search_stats:
        lib.sr_profile_enable(2)
        reg.align(src.view, tgt.view)
        lib.sr_profile_enable(0)
        for name in ("knn_normals", "gicp_pass"):
            st = (ctypes.c_uint64 * 10)()
            lib.sr_profile_read_stats(name.encode(), st)
            q = max(st[0], 1)
            print(f"search stats {name}: queries {st[0]} rings/q {st[1] / q:.2f} "
                  f"cells/q {st[2] / q:.2f} points/q {st[3] / q:.2f} outcomes "
                  f"rejected/block/ring-found/ring-none {list(st[4:8])} max cells {st[8]} "
                  f"max points {st[9]}", file=sys.stderr, flush=True)

    err = None
    golden = os.path.join(ROOT, "tests", "golden", "fp64_poses.json")
    if os.path.exists(golden):
        ref = json.load(open(golden)).get(w.name)
        if ref:
            err = {"rot_rad": synth.rotation_angle_between(res.rotation64, ref["q"]),
                   "trans_m": float(np.linalg.norm(res.translation64 - np.array(ref["t"]))),
                   "vs": "FP64 CPU oracle (tests/golden/fp64_poses.json)"}
    gt = {"rot_rad": synth.rotation_angle_between(res.rotation64, w.gt_q),
          "trans_m": float(np.linalg.norm(res.translation64 - w.gt_t))}

    line = {
        "metric": METRIC,
        "value": args.steps / total,
        "unit": "registrations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 per-point, f64 reduce/solve/pose",
        "data": "synthetic (seeded scene generator, spatialrust_amd/synth.py)",
        "config": {"workload": w.name, "target_points": int(len(w.target)),
                   "source_points": int(len(w.source)), "voxel_leaf": w.voxel_leaf,
                   "k_neighbors": w.k_neighbors,
                   "max_correspondence_distance": w.max_correspondence_distance,
                   "iterations": w.max_iterations, "mode": "plain (one registration per call)",
                   "after_voxel": {"source": res.source_points, "target": res.target_points}},
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                       "p99": float(np.percentile(lat, 99) * 1e3)},
        "error_vs_fp64_oracle": err,
        "error_vs_ground_truth": gt,
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in kern.items()},
        "ms_per_step_eager_timed_launches": total_eager / args.steps * 1e3,
        "roofline": {"kernel": "gicp pass (k_gicp_search_fused + k_gicp_hard + k_gicp_terms)",
                     "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": _traffic("gicp_pass", w.name),
                     "algorithmic_bytes_per_launch": algo_bytes,
                     "avg_launch_us": avg_pass_s * 1e6, "launches": pass_n},
        "cpu_baseline": None,
    }
    if not args.no_batch:
        print("batched vs plain leg ...", file=sys.stderr, flush=True)
        line["batched_vs_plain"] = _batched_vs_plain(args.batch_count)
    if not args.no_serve:
        print("serving leg ...", file=sys.stderr, flush=True)
        line["serving"] = _serving(args.serve_rates, args.serve_seconds, args.serve_max_requests,
                                   args.serve_max_active)
    if not args.no_odometry:
        print("odometry leg ...", file=sys.stderr, flush=True)
        line["odometry"] = _odometry(args.odometry_scans)
    if not args.no_radius:
        print("radius covariance leg ...", file=sys.stderr, flush=True)
        line["radius_covariances"] = _radius_leg(w)
    if not args.no_extra_legs:
        print("host-view and reference-corner legs ...", file=sys.stderr, flush=True)
        line["host_view"] = _host_view_rate(w, cfg, steps=max(3, args.steps // 4))
        line["reference_corner_gpu"] = _reference_corner_gpu()
    if not args.no_cpu_baseline:
        print("timing the CPU oracle (full registration, 1 thread, then all cores) ...",
              file=sys.stderr, flush=True)
        line["cpu_baseline"] = _cpu_baseline(w, min(16, os.cpu_count() or 1))
    line["launch_tuning"] = sr.tune_describe().strip().splitlines()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
