"""C ABI checks that need no GPU: the library loads, exports every function the
public header declares, and the host-side policy contract
(crates/spatialrust-core/src/execution.rs:17-58) holds before any device work."""
import ctypes
import os
import re

import numpy as np
import pytest

from spatialrust_amd import native
import spatialrust_amd as sr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(sr_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_header_symbol():
    lib = native.load()
    names = _header_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
        assert name in native.SIGNATURES, f"{name} has no ctypes signature"


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors agree with the C compiler's layout of the header structs."""
    structs = {"sr_xyz_view": native.XyzView, "sr_gicp_config": native.GicpConfigC,
               "sr_result": native.ResultC, "sr_receipt": native.ReceiptC,
               "sr_policy": native.PolicyC, "sr_server_config": native.ServerConfigC}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sr_gicp.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            a, b, c = line.split()
            got[(a, b)] = int(c)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_config_default_matches_reference_defaults():
    c = native.GicpConfigC()
    native.load().sr_gicp_config_default(ctypes.byref(c))
    # gicp.rs:46-60
    assert c.max_iterations == 50
    assert c.max_correspondence_distance == 1.0
    assert c.k_neighbors == 20
    assert c.epsilon == 1e-3
    assert c.transformation_epsilon == 1e-8
    assert c.fitness_epsilon == 1e-6
    assert c.min_correspondences == 6
    assert list(c.init_q) == [0, 0, 0, 1] and list(c.init_t) == [0, 0, 0]
    py = sr.GicpConfig()._c()
    assert bytes(py) == bytes(c)


def test_gpu_policy_on_cpu_device_is_invalid():
    pts = np.zeros((4, 3), np.float32)
    with pytest.raises(sr.InvalidArgument, match="GPU execution policy cannot target the CPU device"):
        sr.GicpRegistration().align_with_policy(pts, pts, sr.ExecutionPolicy.gpu(sr.DeviceKind.CPU))


def test_cpu_policies_are_left_to_the_caller():
    pts = np.zeros((4, 3), np.float32)
    for pol in (sr.ExecutionPolicy.cpu_single(), sr.ExecutionPolicy.cpu_parallel()):
        with pytest.raises(sr.DeviceUnavailable):
            sr.GicpRegistration().align_with_policy(pts, pts, pol)


def test_wgpu_policy_is_unsupported_here():
    pts = np.zeros((4, 3), np.float32)
    with pytest.raises(sr.InvalidArgument, match="unsupported GPU device kind"):
        sr.GicpRegistration().align_with_policy(pts, pts, sr.ExecutionPolicy.gpu(sr.DeviceKind.WGPU))


@pytest.mark.skipif(native.load().sr_device_available() == 1, reason="a HIP device is present")
def test_without_device_auto_reports_unavailable_and_gpu_errors():
    pts = np.zeros((4, 3), np.float32)
    with pytest.raises(sr.DeviceUnavailable):
        sr.GicpRegistration().align_with_policy(pts, pts, sr.ExecutionPolicy.auto())
    with pytest.raises(sr.InvalidArgument, match="HIP device unavailable"):
        sr.GicpRegistration().align_with_policy(pts, pts, sr.ExecutionPolicy.gpu())


def test_device_kind_is_gpu():
    assert sr.DeviceKind.HIP.is_gpu() and not sr.DeviceKind.CPU.is_gpu()
    assert sr.ExecutionPolicy.auto().allows_fallback()
    assert not sr.ExecutionPolicy.gpu().allows_fallback()


def test_pipeline_transform_point_cloud_matches_oracle(orc):
    """The MVP stage's source synthesis (transform_point_cloud, f32) equals the
    oracle's restatement of Isometry3::transform_point bit for bit."""
    from spatialrust_amd.pipeline import transform_point_cloud
    from spatialrust_amd.registration import Isometry3
    rng = np.random.default_rng(3)
    pts = (rng.standard_normal((5000, 3)) * 20).astype(np.float32)
    for q, t in (((0.0, 0.0, np.sin(0.025), np.cos(0.025)), (0.01, -0.008, 0.012)),
                 ((0.3, -0.2, 0.1, 0.9), (1.5, 2.0, -3.0)), ((0.0, 0.0, 0.0, 0.0), (0.0, 0.0, 1.0))):
        got = transform_point_cloud(pts, Isometry3(q, t))
        want = orc.transform_points(np.array(q, np.float32), np.array(t, np.float32), pts)
        assert np.array_equal(got, want)


def test_pipeline_rejects_non_gicp_methods():
    from spatialrust_amd import pipeline as pl
    from spatialrust_amd.registration import InvalidArgument
    with pytest.raises(InvalidArgument, match="no HIP backend"):
        pl.registration_stage(np.zeros((10, 3), np.float32), pl.MvpIcpConfig())


def test_runtime_options_round_trip_and_errors():
    """sr_option_set / sr_option_get need no device: known names round-trip,
    unknown names and out-of-range values are InvalidArgument."""
    assert sr.get_option("match_reuse") == 1
    sr.set_option("match_reuse", 0)
    try:
        assert sr.get_option("match_reuse") == 0
    finally:
        sr.set_option("match_reuse", 1)
    with pytest.raises(sr.InvalidArgument, match="unknown option"):
        sr.set_option("no_such_option", 1)
    with pytest.raises(sr.InvalidArgument, match="0 or 1"):
        sr.set_option("match_reuse", 2)
    for name, default in (("radius_tiles", 0), ("autotune", 1), ("odometry_incremental", 1),
                          ("pass_graphs", 1)):
        assert sr.get_option(name) == default
        sr.set_option(name, 1 - default)
        try:
            assert sr.get_option(name) == 1 - default
        finally:
            sr.set_option(name, default)
