"""Writes the FP64-oracle poses of the seeded benchmark workloads to
tests/golden/fp64_poses.json (run once; the workloads are deterministic).

    python tests/golden/make_fp64_poses.py

The oracle here is the FP64 restatement of GicpRegistration::align_with_diagnostics
(oracle/oracle.cpp); inputs come from spatialrust_amd.synth with fixed seeds, and
every workload runs its fixed iteration count (transformation_epsilon =
fitness_epsilon = 0).
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from spatialrust_amd import synth  # noqa: E402
from spatialrust_amd.registration import GicpConfig  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fp64_poses.json")


def run(w):
    t0 = time.time()
    sv, _, _ = oracle.voxel_centroid(w.source, w.voxel_leaf)
    tv, _, _ = oracle.voxel_centroid(w.target, w.voxel_leaf)
    cfg = GicpConfig(max_iterations=w.max_iterations, k_neighbors=w.k_neighbors,
                     max_correspondence_distance=w.max_correspondence_distance,
                     transformation_epsilon=0.0, fitness_epsilon=0.0)
    r = oracle.gicp_align(sv, tv, cfg._c(), mode="fp64")
    return {"q": r["q"].tolist(), "t": r["t"].tolist(), "fitness": r["fitness"],
            "iterations": r["iterations"], "source_points": len(sv), "target_points": len(tv),
            "gt_q": w.gt_q.tolist(), "gt_t": w.gt_t.tolist(), "oracle_seconds": time.time() - t0}


def main():
    out = {}
    for w in (synth.small_cube_pair(), synth.large_block_pair()):
        out[w.name] = run(w)
        print(w.name, out[w.name], flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
