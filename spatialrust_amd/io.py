"""PCD and PLY point-cloud files (host side), for feeding real scans to the
device path.

Formats follow the reference's readers and writers:
  PCD v0.7 (crates/spatialrust-io/src/pcd): DATA ascii, binary (point-major
      records) and binary_compressed (two little-endian u32 sizes, then an LZF
      stream of the field-major payload: all values of field 0, then field 1...).
  PLY (crates/spatialrust-io/src/ply): format ascii 1.0 and
      binary_little_endian 1.0, one `vertex` element of scalar properties.

A cloud is returned as a dict of named 1-D numpy columns (x, y, z, intensity,
...); `xyz(cloud)` stacks the positions as an (N, 3) float32 array.
"""
import io
import os
import struct
from typing import Dict, Union

import numpy as np

_PCD_TYPES = {("F", 4): np.float32, ("F", 8): np.float64, ("I", 1): np.int8, ("I", 2): np.int16,
              ("I", 4): np.int32, ("I", 8): np.int64, ("U", 1): np.uint8, ("U", 2): np.uint16,
              ("U", 4): np.uint32, ("U", 8): np.uint64}
_PLY_TYPES = {"char": np.int8, "int8": np.int8, "uchar": np.uint8, "uint8": np.uint8,
              "short": np.int16, "int16": np.int16, "ushort": np.uint16, "uint16": np.uint16,
              "int": np.int32, "int32": np.int32, "uint": np.uint32, "uint32": np.uint32,
              "float": np.float32, "float32": np.float32, "double": np.float64,
              "float64": np.float64}


class PointFileError(ValueError):
    """Malformed or unsupported point-cloud file (SpatialError::Io in the reference)."""


def _read_bytes(src) -> bytes:
    if isinstance(src, (bytes, bytearray)):
        return bytes(src)
    if hasattr(src, "read"):
        return src.read()
    with open(os.fspath(src), "rb") as f:
        return f.read()


def xyz(cloud: Dict[str, np.ndarray]) -> np.ndarray:
    """(N, 3) float32 positions of a cloud read by read_pcd / read_ply."""
    try:
        return np.stack([cloud["x"], cloud["y"], cloud["z"]], 1).astype(np.float32)
    except KeyError as e:
        raise PointFileError(f"cloud has no position field {e}") from None


# ------------------------------------------------------------------- LZF ---

def lzf_decompress(data: bytes, out_len: int) -> bytes:
    """LZF (liblzf) stream decoder: a control byte below 32 starts a literal run
    of ctrl + 1 bytes; otherwise length = ctrl >> 5 (7 means an extra length
    byte follows) and the match is (length + 2) bytes copied from
    ((ctrl & 31) << 8) + next byte + 1 bytes back in the output."""
    out = bytearray(out_len)
    ip = op = 0
    n = len(data)
    while ip < n:
        ctrl = data[ip]
        ip += 1
        if ctrl < 32:
            ln = ctrl + 1
            if ip + ln > n or op + ln > out_len:
                raise PointFileError("truncated LZF literal run in binary_compressed PCD")
            out[op:op + ln] = data[ip:ip + ln]
            ip += ln
            op += ln
            continue
        ln = ctrl >> 5
        back = ((ctrl & 31) << 8) + 1
        if ln == 7:
            if ip >= n:
                raise PointFileError("truncated LZF length in binary_compressed PCD")
            ln += data[ip]
            ip += 1
        if ip >= n:
            raise PointFileError("truncated LZF back-reference in binary_compressed PCD")
        back += data[ip]
        ip += 1
        ln += 2
        if back > op or op + ln > out_len:
            raise PointFileError("invalid LZF back-reference in binary_compressed PCD")
        for k in range(ln):  # byte by byte: the source may overlap the output
            out[op + k] = out[op - back + k]
        op += ln
    if op != out_len:
        raise PointFileError(f"LZF payload size mismatch: expected {out_len}, decoded {op}")
    return bytes(out)


def lzf_compress(data: bytes) -> bytes:
    """A valid LZF stream for `data`: greedy matches of 3+ bytes found through a
    hash of the next three bytes (offsets up to 8192), literal runs otherwise."""
    out = bytearray()
    lit = bytearray()
    table = {}
    i, n = 0, len(data)

    def flush():
        for s in range(0, len(lit), 32):
            chunk = lit[s:s + 32]
            out.append(len(chunk) - 1)
            out.extend(chunk)
        lit.clear()

    while i < n:
        key = data[i:i + 3]
        j = table.get(key) if len(key) == 3 else None
        if len(key) == 3:
            table[key] = i
        if j is not None and 0 < i - j <= 8192:
            ln = 3
            while i + ln < n and ln < 264 and data[j + ln] == data[i + ln]:
                ln += 1
            flush()
            back = i - j - 1
            code = ln - 2
            if code < 7:
                out.append((code << 5) | (back >> 8))
            else:
                out.append((7 << 5) | (back >> 8))
                out.append(code - 7)
            out.append(back & 0xFF)
            i += ln
        else:
            lit.append(data[i])
            i += 1
    flush()
    return bytes(out)


# ------------------------------------------------------------------- PCD ---

def _pcd_header(buf: bytes):
    pos = 0
    hdr = {}
    while True:
        end = buf.find(b"\n", pos)
        if end < 0:
            raise PointFileError("PCD header missing DATA")
        line = buf[pos:end].decode("ascii", "replace").strip()
        pos = end + 1
        if not line or line.startswith("#"):
            continue
        key, _, rest = line.partition(" ")
        hdr[key.upper()] = rest.split()
        if key.upper() == "DATA":
            break
    for need in ("FIELDS", "SIZE", "TYPE"):
        if need not in hdr:
            raise PointFileError(f"PCD header missing {need}")
    fields = hdr["FIELDS"]
    count = [int(c) for c in hdr.get("COUNT", ["1"] * len(fields))]
    size = [int(s) for s in hdr["SIZE"]]
    types = hdr["TYPE"]
    if not (len(fields) == len(size) == len(types) == len(count)):
        raise PointFileError("PCD FIELDS/SIZE/TYPE/COUNT lengths differ")
    dt = []
    for f, s, t in zip(fields, size, types):
        if (t, s) not in _PCD_TYPES:
            raise PointFileError(f"unsupported PCD field type {t}{s}")
        dt.append(_PCD_TYPES[(t, s)])
    width = int(hdr.get("WIDTH", ["0"])[0])
    height = int(hdr.get("HEIGHT", ["1"])[0])
    points = int(hdr.get("POINTS", [str(width * height)])[0])
    mode = hdr["DATA"][0] if hdr["DATA"] else ""
    if mode not in ("ascii", "binary", "binary_compressed"):
        raise PointFileError(f"unsupported DATA mode `{mode}`")
    return fields, count, dt, points, mode, pos


def read_pcd(src) -> Dict[str, np.ndarray]:
    """PCD v0.7 file (path, bytes or binary stream) -> dict of columns. Fields
    with COUNT > 1 become (N, COUNT) arrays."""
    buf = _read_bytes(src)
    fields, count, dt, n, mode, pos = _pcd_header(buf)
    if mode == "ascii":
        rows = [ln.split() for ln in buf[pos:].decode("ascii", "replace").splitlines() if ln.strip()]
        if len(rows) < n:
            raise PointFileError(f"PCD has {len(rows)} ascii rows, POINTS says {n}")
        vals = np.array(rows[:n], dtype=np.float64).reshape(n, -1)
        out, col = {}, 0
        for f, c, d in zip(fields, count, dt):
            v = vals[:, col:col + c].astype(d)
            out[f] = v[:, 0] if c == 1 else v
            col += c
        return out
    record = np.dtype([(f, np.dtype(d).newbyteorder("<"), (c,)) for f, c, d in zip(fields, count, dt)])
    if mode == "binary":
        need = record.itemsize * n
        if len(buf) - pos < need:
            raise PointFileError(f"binary payload size mismatch: expected {need}, found {len(buf) - pos}")
        arr = np.frombuffer(buf, dtype=record, count=n, offset=pos)
        return {f: (arr[f][:, 0] if c == 1 else arr[f]).astype(d) for f, c, d in zip(fields, count, dt)}
    if len(buf) - pos < 8:
        raise PointFileError("truncated binary_compressed PCD sizes")
    csize, usize = struct.unpack_from("<II", buf, pos)
    payload = lzf_decompress(buf[pos + 8:pos + 8 + csize], usize)
    if usize != record.itemsize * n:
        raise PointFileError(f"binary_compressed payload size mismatch: expected "
                             f"{record.itemsize * n}, found {usize}")
    out, off = {}, 0
    for f, c, d in zip(fields, count, dt):  # field-major
        nb = np.dtype(d).itemsize * c * n
        v = np.frombuffer(payload, dtype=np.dtype(d).newbyteorder("<"), count=c * n, offset=off)
        off += nb
        v = v.astype(d).reshape(n, c)
        out[f] = v[:, 0] if c == 1 else v
    return out


def write_pcd(dst, points, fmt: str = "binary", extra: Dict[str, np.ndarray] = None):
    """Writes xyz (N, 3) float32 (plus optional float32 columns) as PCD v0.7 with
    DATA `fmt` in {ascii, binary, binary_compressed}. dst: path or binary stream."""
    p = np.asarray(points, np.float32).reshape(-1, 3)
    cols = {"x": p[:, 0], "y": p[:, 1], "z": p[:, 2]}
    for k, v in (extra or {}).items():
        cols[k] = np.asarray(v, np.float32)
    n = len(p)
    names = list(cols)
    head = ("# .PCD v0.7 - Point Cloud Data file format\nVERSION 0.7\n"
            f"FIELDS {' '.join(names)}\nSIZE {' '.join(['4'] * len(names))}\n"
            f"TYPE {' '.join(['F'] * len(names))}\nCOUNT {' '.join(['1'] * len(names))}\n"
            f"WIDTH {n}\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS {n}\nDATA {fmt}\n").encode()
    if fmt == "ascii":
        body = "".join(" ".join(repr(float(cols[k][i])) for k in names) + "\n" for i in range(n)).encode()
    elif fmt == "binary":
        body = np.stack([cols[k] for k in names], 1).astype("<f4").tobytes()
    elif fmt == "binary_compressed":
        raw = b"".join(cols[k].astype("<f4").tobytes() for k in names)
        comp = lzf_compress(raw)
        body = struct.pack("<II", len(comp), len(raw)) + comp
    else:
        raise PointFileError(f"unsupported DATA mode `{fmt}`")
    _write(dst, head + body)


# ------------------------------------------------------------------- PLY ---

def read_ply(src) -> Dict[str, np.ndarray]:
    """PLY (ascii 1.0 / binary_little_endian 1.0) -> dict of `vertex` columns."""
    buf = _read_bytes(src)
    if not buf.startswith(b"ply"):
        raise PointFileError("missing PLY magic")
    end = buf.find(b"end_header")
    if end < 0:
        raise PointFileError("PLY header missing end_header")
    body = buf.index(b"\n", end) + 1
    fmt, elems, cur = None, [], None
    for line in buf[:end].decode("ascii", "replace").splitlines()[1:]:
        tok = line.split()
        if not tok or tok[0] in ("comment", "obj_info"):
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            cur = {"name": tok[1], "count": int(tok[2]), "props": []}
            elems.append(cur)
        elif tok[0] == "property":
            if cur is None:
                raise PointFileError("PLY property before element")
            if tok[1] == "list":
                raise PointFileError("PLY list properties are not supported")
            if tok[1] not in _PLY_TYPES:
                raise PointFileError(f"unsupported PLY property type `{tok[1]}`")
            cur["props"].append((tok[2], _PLY_TYPES[tok[1]]))
    if fmt not in ("ascii", "binary_little_endian"):
        raise PointFileError(f"unsupported PLY format `{fmt}`")
    if not elems or elems[0]["name"] != "vertex":
        raise PointFileError("PLY file must start with a vertex element")
    v = elems[0]
    n = v["count"]
    if fmt == "ascii":
        rows = [ln.split() for ln in buf[body:].decode("ascii", "replace").splitlines()[:n]]
        vals = np.array(rows, dtype=np.float64).reshape(n, len(v["props"]))
        return {name: vals[:, i].astype(d) for i, (name, d) in enumerate(v["props"])}
    rec = np.dtype([(name, np.dtype(d).newbyteorder("<")) for name, d in v["props"]])
    if len(buf) - body < rec.itemsize * n:
        raise PointFileError("truncated binary PLY vertex data")
    arr = np.frombuffer(buf, dtype=rec, count=n, offset=body)
    return {name: arr[name].astype(d) for name, d in v["props"]}


def write_ply(dst, points, fmt: str = "binary_little_endian", extra: Dict[str, np.ndarray] = None):
    """Writes xyz (N, 3) float32 (plus optional float32 columns) as one PLY vertex
    element, `fmt` in {ascii, binary_little_endian}."""
    p = np.asarray(points, np.float32).reshape(-1, 3)
    cols = {"x": p[:, 0], "y": p[:, 1], "z": p[:, 2]}
    for k, val in (extra or {}).items():
        cols[k] = np.asarray(val, np.float32)
    n = len(p)
    if fmt not in ("ascii", "binary_little_endian"):
        raise PointFileError(f"unsupported PLY format `{fmt}`")
    head = (f"ply\nformat {fmt} 1.0\nelement vertex {n}\n" +
            "".join(f"property float {k}\n" for k in cols) + "end_header\n").encode()
    if fmt == "ascii":
        body = "".join(" ".join(repr(float(cols[k][i])) for k in cols) + "\n" for i in range(n)).encode()
    else:
        body = np.stack(list(cols.values()), 1).astype("<f4").tobytes()
    _write(dst, head + body)


def _write(dst, data: bytes):
    if hasattr(dst, "write"):
        dst.write(data)
    else:
        with open(os.fspath(dst), "wb") as f:
            f.write(data)
