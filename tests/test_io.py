"""PCD / PLY files (host side). Cases follow the reference's reader tests
(crates/spatialrust-io/src/pcd/reader.rs:588-715, ply/reader.rs:420-499)."""
import io
import struct

import numpy as np
import pytest

from spatialrust_amd import io as sio

XYZ_PCD = b"""# .PCD v0.7 - Point Cloud Data file format
VERSION 0.7
FIELDS x y z
SIZE 4 4 4
TYPE F F F
COUNT 1 1 1
WIDTH 3
HEIGHT 1
VIEWPOINT 0 0 0 1 0 0 0
POINTS 3
DATA ascii
0.0 0.0 0.0
1.0 0.0 0.0
0.0 1.0 0.0
"""


def test_pcd_ascii_xyz_and_intensity():
    c = sio.read_pcd(XYZ_PCD)
    assert np.array_equal(c["x"], [0, 1, 0]) and np.array_equal(c["y"], [0, 0, 1])
    assert np.array_equal(c["z"], [0, 0, 0])
    xyzi = XYZ_PCD.replace(b"FIELDS x y z", b"FIELDS x y z intensity").replace(
        b"SIZE 4 4 4", b"SIZE 4 4 4 4").replace(b"TYPE F F F", b"TYPE F F F F").replace(
        b"COUNT 1 1 1", b"COUNT 1 1 1 1").replace(b"0.0 0.0 0.0\n1.0 0.0 0.0\n0.0 1.0 0.0\n",
                                                  b"0.0 0.0 0.0 0.5\n1.0 0.0 0.0 0.8\n0.0 1.0 0.0 0.1\n")
    assert np.array_equal(sio.read_pcd(xyzi)["intensity"], np.float32([0.5, 0.8, 0.1]))


def test_pcd_binary_compressed_literal_stream():
    # field-major payload x0 x1 y0 y1 z0 z1 as one LZF literal run
    raw = np.float32([1, 2, 3, 4, 5, 6]).tobytes()
    comp = bytes([len(raw) - 1]) + raw
    head = XYZ_PCD.split(b"DATA")[0].replace(b"WIDTH 3", b"WIDTH 2").replace(b"POINTS 3", b"POINTS 2")
    data = head + b"DATA binary_compressed\n" + struct.pack("<II", len(comp), len(raw)) + comp
    c = sio.read_pcd(data)
    assert np.array_equal(c["x"], [1, 2]) and np.array_equal(c["y"], [3, 4])
    assert np.array_equal(c["z"], [5, 6])


@pytest.mark.parametrize("fmt", ["ascii", "binary", "binary_compressed"])
def test_pcd_roundtrip(fmt, tmp_path):
    rng = np.random.default_rng(1)
    pts = (rng.standard_normal((2000, 3)) * 10).astype(np.float32)
    pts[:500] = pts[500:1000]  # repeats give the compressor back-references
    path = tmp_path / f"c_{fmt}.pcd"
    sio.write_pcd(path, pts, fmt, extra={"intensity": pts[:, 0] * 2})
    c = sio.read_pcd(path)
    assert np.array_equal(sio.xyz(c), pts)
    assert np.array_equal(c["intensity"], pts[:, 0] * 2)


def test_lzf_roundtrip_and_overlapping_match():
    rng = np.random.default_rng(2)
    for data in (b"", b"a", b"abcabcabcabcabcabc" * 50, bytes(rng.integers(0, 4, 5000, dtype=np.uint8)),
                 bytes(rng.integers(0, 256, 3000, dtype=np.uint8))):
        assert sio.lzf_decompress(sio.lzf_compress(data), len(data)) == data
    # literal "ab", then a match of length 6 from 2 back (overlapping copy)
    stream = bytes([1]) + b"ab" + bytes([(4 << 5) | 0, 1])
    assert sio.lzf_decompress(stream, 8) == b"abababab"
    with pytest.raises(sio.PointFileError):
        sio.lzf_decompress(bytes([(1 << 5) | 0, 5]), 3)  # reference before the start


def test_ply_ascii_and_binary():
    ply = b"ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n" \
          b"property float z\nproperty float intensity\nend_header\n" \
          b"0 0 0 0.5\n1 0 0 0.8\n0 1 0 0.1\n"
    c = sio.read_ply(ply)
    assert np.array_equal(c["x"], [0, 1, 0]) and np.array_equal(c["y"], [0, 0, 1])
    assert np.array_equal(c["intensity"], np.float32([0.5, 0.8, 0.1]))
    pts = np.float32([[0.5, 1.5, 2.5], [1, 2, 3]])
    for fmt in ("ascii", "binary_little_endian"):
        buf = io.BytesIO()
        sio.write_ply(buf, pts, fmt)
        assert np.array_equal(sio.xyz(sio.read_ply(buf.getvalue())), pts)


def test_errors():
    with pytest.raises(sio.PointFileError, match="DATA mode"):
        sio.read_pcd(XYZ_PCD.replace(b"DATA ascii", b"DATA zip"))
    with pytest.raises(sio.PointFileError, match="PLY magic"):
        sio.read_ply(b"plx\n")
    with pytest.raises(sio.PointFileError, match="format"):
        sio.read_ply(b"ply\nformat binary_big_endian 1.0\nelement vertex 0\nend_header\n")
