"""The allocation-free GICP entry point (sr_gicp_align_noalloc) against
sr_gicp_align, through the C ABI (gpu).

Bar: bit-identical results (FP64 pose, fitness, iterations, convergence flag,
point and correspondence counts) on two calls in a row, the second with
smaller clouds, so the second call runs over scratch the first one filled.
"""
import ctypes

import numpy as np
import pytest

import spatialrust_amd as sr
from spatialrust_amd import devmem, native, synth

pytestmark = pytest.mark.gpu


def _call(fn_name, sv, tv, cfg):
    lib = native.load()
    out = native.ResultC()
    c = cfg._c()
    if fn_name == "sr_gicp_align":
        pol = native.PolicyC(native.SR_GPU, native.SR_DEV_HIP)
        code = lib.sr_gicp_align(ctypes.byref(sv), ctypes.byref(tv), ctypes.byref(c), pol,
                                 ctypes.byref(out), None)
    else:
        code = lib.sr_gicp_align_noalloc(ctypes.byref(sv), ctypes.byref(tv), ctypes.byref(c),
                                         ctypes.byref(out))
    return code, out


def _fields(r):
    return (tuple(r.q64), tuple(r.t64), r.fitness, r.iterations, r.converged, r.source_points,
            r.target_points, r.correspondences, tuple(r.q), tuple(r.t))


def _host_view(pts):
    cols = [np.ascontiguousarray(pts[:, a], dtype=np.float32) for a in range(3)]
    v = native.XyzView(cols[0].ctypes.data, cols[1].ctypes.data, cols[2].ctypes.data, len(pts),
                       native.SR_MEM_HOST)
    return v, cols


def test_noalloc_equals_align_on_two_calls(hip):
    lib = native.load()
    scratch = ctypes.c_int64(0)
    assert lib.sr_option_get(b"noalloc_scratch_bytes", ctypes.byref(scratch)) == 0
    cfg = sr.GicpConfig(max_iterations=30, voxel_leaf=0.1, transformation_epsilon=0.0,
                        fitness_epsilon=0.0)
    # call 1: small-cube pair (50k + 50k points), device-resident columns
    w1 = synth.small_cube_pair()
    s1, t1 = devmem.DeviceCloud(w1.source), devmem.DeviceCloud(w1.target)
    try:
        print("noalloc call 1: 50k + 50k points, device views", flush=True)
        code_a, ref1 = _call("sr_gicp_align", s1.view, t1.view, cfg)
        code_b, got1 = _call("sr_gicp_align_noalloc", s1.view, t1.view, cfg)
    finally:
        s1.close()
        t1.close()
    assert code_a == 0, native.last_error()
    assert code_b == 0, native.last_error()
    assert _fields(got1) == _fields(ref1)
    peak1 = ctypes.c_int64(0)
    assert lib.sr_option_get(b"noalloc_peak_bytes", ctypes.byref(peak1)) == 0
    assert 0 < peak1.value <= scratch.value
    # call 2: a smaller pair (15k + 20k), host columns, default epsilons
    w2 = synth.small_cube_pair(instance=3, n_source=15000, n_target=20000)
    cfg2 = sr.GicpConfig(max_iterations=30, voxel_leaf=0.1)
    sv, ks = _host_view(w2.source)
    tv, kt = _host_view(w2.target)
    print("noalloc call 2: 15k + 20k points, host views", flush=True)
    code_a, ref2 = _call("sr_gicp_align", sv, tv, cfg2)
    code_b, got2 = _call("sr_gicp_align_noalloc", sv, tv, cfg2)
    assert code_a == 0 and code_b == 0, native.last_error()
    assert _fields(got2) == _fields(ref2)
    peak2 = ctypes.c_int64(0)
    assert lib.sr_option_get(b"noalloc_peak_bytes", ctypes.byref(peak2)) == 0
    assert 0 < peak2.value < peak1.value
    print(f"noalloc peaks: {peak1.value} and {peak2.value} of {scratch.value} bytes", flush=True)
    del ks, kt


def test_noalloc_errors(hip):
    """Clouds above the stated maximum are refused before any work; the
    reference's GICP errors come through unchanged."""
    lib = native.load()
    big = np.zeros((2_000_001, 3), np.float32)
    small = synth.small_cube_pair(n_source=3000, n_target=3000)
    bv, kb = _host_view(big)
    sv, ks = _host_view(small.source)
    out = native.ResultC()
    cfg = sr.GicpConfig()._c()
    assert lib.sr_gicp_align_noalloc(ctypes.byref(bv), ctypes.byref(sv), ctypes.byref(cfg),
                                     ctypes.byref(out)) == native.SR_ERR_INVALID_ARGUMENT
    assert "at most 2000000 points" in native.last_error()
    far, kf = _host_view(small.source + np.float32(500.0))
    tv, kt = _host_view(small.target)
    assert lib.sr_gicp_align_noalloc(ctypes.byref(far), ctypes.byref(tv), ctypes.byref(cfg),
                                     ctypes.byref(out)) == native.SR_ERR_INVALID_ARGUMENT
    assert native.last_error() == "GICP found only 0 correspondences, minimum is 6"
    del kb, ks, kf, kt
