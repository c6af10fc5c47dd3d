"""Binary scene and query dumps, the files the reference's benches exchange with
external engines so that both sides see byte-identical FP32 inputs.

QRYD (bench/gpu_query_bench.cpp:232-259): little-endian u32 magic 0x44595251
("QRYD" in file order), u32 version 1, u64 triangle count, u64 point count, then
the triangles as 9 floats (a, b, c) and the query points as 3 floats.

QRCD (bench/gpu_cloud_bench.cpp:323-343): the 4 bytes "QRCD", u32 version 1,
u64 point count, u64 query count, then the points and the queries as 3 floats.

Both headers are 24 bytes.  Readers return float32 arrays (triangles (n, 3, 3),
points and queries (n, 3)) viewing a copy of the file body; they raise ValueError
on a wrong magic, version or length.
"""
from __future__ import annotations

import struct

import numpy as np

QRYD_MAGIC = b"QRYD"
QRCD_MAGIC = b"QRCD"
VERSION = 1
HEADER = struct.Struct("<4sIQQ")  # magic, version, count a, count b


def _write(path, magic: bytes, a: np.ndarray, b: np.ndarray, a_floats: int) -> None:
    a = np.ascontiguousarray(a, dtype="<f4").reshape(-1, a_floats)
    b = np.ascontiguousarray(b, dtype="<f4").reshape(-1, 3)
    with open(path, "wb") as f:
        f.write(HEADER.pack(magic, VERSION, a.shape[0], b.shape[0]))
        f.write(a.tobytes())
        f.write(b.tobytes())


def _read(path, magic: bytes, a_floats: int):
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < HEADER.size:
        raise ValueError(f"{path}: shorter than the {HEADER.size}-byte header")
    m, version, na, nb = HEADER.unpack_from(raw, 0)
    if m != magic:
        raise ValueError(f"{path}: magic {m!r}, expected {magic!r}")
    if version != VERSION:
        raise ValueError(f"{path}: unsupported version {version}")
    want = HEADER.size + 4 * (a_floats * na + 3 * nb)
    if len(raw) != want:
        raise ValueError(f"{path}: {len(raw)} bytes, the header announces {want}")
    body = np.frombuffer(raw, dtype="<f4", offset=HEADER.size).astype(np.float32)
    return body[: a_floats * na].reshape(na, a_floats), body[a_floats * na:].reshape(nb, 3)


def write_qryd(path, triangles, points) -> None:
    """Triangles (n, 3, 3) or (n, 9) and query points (m, 3) as a QRYD file."""
    _write(path, QRYD_MAGIC, triangles, points, 9)


def read_qryd(path):
    """(triangles (n, 3, 3) float32, points (m, 3) float32) of a QRYD file."""
    tris, pts = _read(path, QRYD_MAGIC, 9)
    return tris.reshape(-1, 3, 3), pts


def write_qrcd(path, points, queries) -> None:
    """A point cloud (n, 3) and query points (m, 3) as a QRCD file."""
    _write(path, QRCD_MAGIC, points, queries, 3)


def read_qrcd(path):
    """(points (n, 3) float32, queries (m, 3) float32) of a QRCD file."""
    return _read(path, QRCD_MAGIC, 3)
