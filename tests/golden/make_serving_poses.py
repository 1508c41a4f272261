"""Writes FP64-oracle poses of three requests of the serving workload (the
requests against the smallest map, with its three smallest sources) to
tests/golden/serving_poses.json.

    python tests/golden/make_serving_poses.py

Config of the serving workload: leaf 0.2, k 15, d_max 2.0, at most 40
iterations, converged when the pose update is below 1e-6 (fitness criterion
off). The oracle is the FP64 restatement of
GicpRegistration::align_with_diagnostics (oracle/oracle.cpp); inputs come from
spatialrust_amd.synth.serving_pool / serving_map / serving_source.
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from spatialrust_amd import synth  # noqa: E402
from spatialrust_amd.registration import GicpConfig  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "serving_poses.json")


def main():
    pool = synth.serving_pool()
    m = int(np.argmin(pool.map_sizes))
    js = [int(j) for j in np.argsort(pool.source_sizes[m], kind="stable")[:3]]
    tgt = synth.serving_map(pool, m)
    tv, _, _ = oracle.voxel_centroid(tgt, 0.2)
    cfg = GicpConfig(max_iterations=40, k_neighbors=15, max_correspondence_distance=2.0,
                     transformation_epsilon=1e-6, fitness_epsilon=0.0)
    out = {"config": "leaf 0.2, k 15, d_max 2.0, 40 iterations, transformation_epsilon 1e-6",
           "requests": {}}
    for j in js:
        t0 = time.time()
        w = synth.serving_source(pool, m, j)
        sv, _, _ = oracle.voxel_centroid(w.source, 0.2)
        r = oracle.gicp_align(sv, tv, cfg._c(), mode="fp64")
        out["requests"][f"{m}_{j}"] = {
            "q": r["q"].tolist(), "t": r["t"].tolist(), "fitness": r["fitness"],
            "iterations": r["iterations"], "converged": r["converged"],
            "raw_source_points": len(w.source), "source_points": len(sv),
            "target_points": len(tv), "oracle_seconds": time.time() - t0}
        print(m, j, out["requests"][f"{m}_{j}"], flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
