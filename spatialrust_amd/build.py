"""Builds libsr_gicp.so (HIP, gfx950) in-tree with hipcc.

    python -m spatialrust_amd.build

Every .hip file under csrc/ is compiled to an object in parallel, then linked
into spatialrust_amd/lib/libsr_gicp.so. Flags:
  -ffp-contract=off   no implicit FMA: the voxel keys, squared distances and
                      radius covariances are meant to round exactly like the
                      reference's f32 code; the GICP kernel writes its fused
                      operations explicitly where it wants them.
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keep the compiler from emitting packed FP32 VALU ops.
  -mllvm -amdgpu-kernarg-preload-count=16
                      the first kernel arguments arrive in SGPRs at wave
                      launch instead of through a load from the argument
                      segment: every kernel's chain of dependent loads starts
                      one round trip earlier (the pass stages are launched
                      some 200 times per registration).
"""
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "lib", "obj")
LIB = os.path.join(LIBDIR, "libsr_gicp.so")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-mllvm", "-amdgpu-kernarg-preload-count=16",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def _hipcc():
    for cand in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(cand) or cand == "hipcc":
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src, verbose):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [_hipcc(), *COMMON, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(line for line in r.stderr.splitlines()
                     if "packed-fp32-ops" not in line and line.strip())
    return obj, warn


def build(verbose: bool = False, jobs: int = 8) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(jobs, len(srcs))) as ex:
        for obj, warn in ex.map(lambda s: _compile(s, verbose), srcs):
            objs.append(obj)
            if warn and verbose:
                print(warn, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
