"""Inputs of the reference's own tests and benches, rebuilt in numpy (f32).

Each builder cites the reference test it reproduces; values are computed with
the same f32 operations so the points are bit-identical.
"""
import numpy as np

F = np.float32


def lcg_points(seed: int, count: int):
    """LCG cloud of kdtree.rs:711-724 (state = state*1664525 + 1013904223,
    value = (state as f32 / u32::MAX as f32) * 20 - 10), then `count`
    further draws are returned as a generator for queries."""
    state = seed & 0xFFFFFFFF

    def nxt():
        nonlocal state
        state = (state * 1664525 + 1013904223) & 0xFFFFFFFF
        return F(F(F(state) / F(4294967295.0)) * F(20.0)) - F(10.0)

    pts = np.array([[nxt(), nxt(), nxt()] for _ in range(count)], dtype=F)
    return pts, nxt


def sample_cloud():
    """kdtree.rs:623-629."""
    return np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 1, 0], [4, 2, 0], [5, 0, 5]], dtype=F)


def box_corner():
    """Three perpendicular faces, 300 points (gicp.rs:452-463)."""
    pts = []
    for i in range(10):
        for j in range(10):
            a = F(F(i) * F(0.08))
            b = F(F(j) * F(0.08))
            pts.append([a, b, F(0.0)])
            pts.append([a, F(0.0), F(b + F(0.04))])
            pts.append([F(0.0), F(a + F(0.04)), F(b + F(0.04))])
    return np.array(pts, dtype=F)


def bench_corner(side: int = 50):
    """Box corner of the registration bench, 3*side^2 points
    (crates/spatialrust-registration/benches/registration.rs:23-45)."""
    step = F(F(1.5) / F(side))
    pts = []
    for i in range(side):
        for j in range(side):
            a = F(F(i) * step)
            b = F(F(j) * step)
            pts.append([a, b, F(0.0)])
            pts.append([a, F(0.0), F(b + F(0.02))])
            pts.append([F(0.0), F(a + F(0.02)), F(b + F(0.02))])
    return np.array(pts, dtype=F)


def sphere_surface(n=300, radius=1.0):
    """Fibonacci sphere like test_bindings.py's sphere_surface helper (300 points, r=1)."""
    i = np.arange(n, dtype=np.float64) + 0.5
    phi = np.arccos(1.0 - 2.0 * i / n)
    theta = np.pi * (1.0 + 5.0 ** 0.5) * i
    pts = np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], 1)
    return (pts * radius).astype(F)


def bench_corner_pair(side: int = 50, yaw: float = 0.03, t=(0.02, -0.015, 0.01)):
    """The registration bench's case (benches/registration.rs:23-63, 81-86):
    target = bench_corner(side), source = target moved by a yaw of `yaw` rad
    about z and translation t, in f32. Returns (target, source, q, t)."""
    tgt = bench_corner(side)
    h = F(F(yaw) * F(0.5))
    q = np.array([0.0, 0.0, np.sin(h), np.cos(h)], dtype=F)
    tv = np.array(t, dtype=F)
    x, y, z, w = (np.float64(v) for v in q)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    src = (tgt.astype(np.float64) @ R.T + tv).astype(F)
    return tgt, src, q, tv


def corner_probe_error(q, t, mq, mt, probe=(0.4, 0.5, 0.3)):
    """|T_est ∘ T_mis (p) - p| at the bench's probe point (FP64): how far the
    estimated transform is from undoing the misalignment."""
    def mat(qq):
        x, y, z, w = (float(v) for v in qq)
        n = (x * x + y * y + z * z + w * w) ** 0.5
        x, y, z, w = x / n, y / n, z / n, w / n
        return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                         [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                         [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    p = np.asarray(probe, np.float64)
    moved = mat(mq) @ p + np.asarray(mt, np.float64)
    back = mat(q) @ moved + np.asarray(t, np.float64)
    return float(np.linalg.norm(back - p))
