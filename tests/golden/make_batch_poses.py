"""Writes the FP64-oracle poses of eight instances of the 256-pair batch
workload (small-cube generator, source sizes log-uniform in [5k, 120k], target
1.5x; leaf 0.1, k 20, d_max 1.0, 30 fixed iterations) to
tests/golden/batch_poses.json.

    python tests/golden/make_batch_poses.py

The instances are the two smallest, the two largest and four in between, so
the golden set spans the batch's size range. The oracle is the FP64
restatement of GicpRegistration::align_with_diagnostics (oracle/oracle.cpp);
the workload is deterministic (spatialrust_amd.synth.batch_pairs, seed offset 0).
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

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch_poses.json")
COUNT = 256


def pick(sizes):
    order = np.argsort(sizes, kind="stable")
    mid = [order[int(f * (len(order) - 1))] for f in (0.3, 0.5, 0.7, 0.85)]
    return sorted(set(int(i) for i in [order[0], order[1], *mid, order[-2], order[-1]]))


def main():
    pairs = synth.batch_pairs(count=COUNT)
    sizes = np.array([len(p.source) for p in pairs])
    cfg = GicpConfig(max_iterations=30, k_neighbors=20, max_correspondence_distance=1.0,
                     transformation_epsilon=0.0, fitness_epsilon=0.0)
    out = {"count": COUNT, "voxel_leaf": 0.1, "instances": {}}
    for i in pick(sizes):
        w = pairs[i]
        t0 = time.time()
        sv, _, _ = oracle.voxel_centroid(w.source, w.voxel_leaf)
        tv, _, _ = oracle.voxel_centroid(w.target, w.voxel_leaf)
        r = oracle.gicp_align(sv, tv, cfg._c(), mode="fp64")
        out["instances"][str(i)] = {
            "q": r["q"].tolist(), "t": r["t"].tolist(), "fitness": r["fitness"],
            "iterations": r["iterations"], "raw_source_points": len(w.source),
            "source_points": len(sv), "target_points": len(tv),
            "oracle_seconds": time.time() - t0}
        print(i, out["instances"][str(i)], flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
