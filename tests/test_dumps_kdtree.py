"""QRYD / QRCD dump files and the scipy cKDTree witness for point queries.

CPU: the byte layout of both formats (bench/gpu_query_bench.cpp:232-259,
bench/gpu_cloud_bench.cpp:323-343), round trips, refusal of damaged files, and the
oracle's point kNN / radius against cKDTree (oracle/kdwitness.py) on a
sphere-shell + cube cloud -- an independent CPU witness for the point configs.
GPU: the reference's own benches (built against this library) write the dumps; the
GPU answers on the dumped inputs equal the witness and the oracle bit for bit.
"""
from __future__ import annotations

import os
import struct
import subprocess

import numpy as np
import pytest

import oracle as O
from oracle import kdwitness as W
from quarry import dumps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "oracle", "_ref", "dropin")


def _cloud(n: int, seed: int) -> np.ndarray:
    """Half on a radius-16 sphere shell, half uniform in a cube of side 64 (the C3
    point cloud's recipe at a smaller count)."""
    rng = np.random.default_rng(seed)
    h = n // 2
    v = rng.normal(size=(h, 3))
    shell = 16.0 * v / np.linalg.norm(v, axis=1, keepdims=True)
    cube = rng.uniform(-32.0, 32.0, (n - h, 3))
    return np.concatenate([shell, cube]).astype(np.float32)


def test_qryd_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    tris = rng.normal(size=(5, 3, 3)).astype(np.float32)
    pts = rng.normal(size=(4, 3)).astype(np.float32)
    p = tmp_path / "scene.qryd"
    dumps.write_qryd(p, tris, pts)
    raw = p.read_bytes()
    assert raw[:4] == b"QRYD"
    # the reference writes the magic as the little-endian u32 0x44595251
    assert struct.unpack_from("<IIQQ", raw, 0) == (0x44595251, 1, 5, 4)
    assert len(raw) == 24 + 4 * (5 * 9 + 4 * 3)
    assert np.array_equal(np.frombuffer(raw, "<f4", 9, 24), tris[0].reshape(-1))
    t2, p2 = dumps.read_qryd(p)
    assert t2.shape == (5, 3, 3) and p2.shape == (4, 3)
    assert np.array_equal(t2.view(np.uint32), tris.view(np.uint32))
    assert np.array_equal(p2.view(np.uint32), pts.view(np.uint32))


def test_qrcd_layout_and_round_trip(tmp_path):
    pts = _cloud(1000, 2)
    q = _cloud(10, 3)
    p = tmp_path / "cloud.qrcd"
    dumps.write_qrcd(p, pts, q)
    raw = p.read_bytes()
    assert raw[:4] == b"QRCD" and struct.unpack_from("<IQQ", raw, 4) == (1, 1000, 10)
    assert len(raw) == 24 + 4 * 3 * 1010
    a, b = dumps.read_qrcd(p)
    assert np.array_equal(a.view(np.uint32), pts.view(np.uint32))
    assert np.array_equal(b.view(np.uint32), q.view(np.uint32))


def test_dump_readers_refuse_damaged_files(tmp_path):
    p = tmp_path / "x.bin"
    dumps.write_qrcd(p, _cloud(10, 4), _cloud(3, 5))
    raw = p.read_bytes()
    for bad in (raw[:-4], raw + b"\0\0\0\0", b"QRYD" + raw[4:], raw[:4] + struct.pack("<I", 2) + raw[8:],
                raw[:10]):
        p.write_bytes(bad)
        with pytest.raises(ValueError):
            dumps.read_qrcd(p)
    p.write_bytes(raw)
    with pytest.raises(ValueError):
        dumps.read_qryd(p)  # a QRCD file is not a QRYD file


def test_witness_guard_on_duplicates():
    """Exact duplicates tie at every rank; the witness keeps (dist2, index) order."""
    pts = np.repeat(_cloud(50, 6), 3, axis=0)
    q = _cloud(20, 7)
    d, i = W.knn(pts, q, 6, extra=12)
    exp = O.knn_brute(1, pts, q, 6, threads=4)
    assert np.array_equal(d.view(np.uint32), exp[:, :, 0])
    assert np.array_equal(i, exp[:, :, 1].view(np.int32))


@pytest.mark.parametrize("k", [1, 8, 32])
def test_oracle_points_knn_match_ckdtree(k):
    pts = _cloud(1 << 15, 8)
    q = np.random.default_rng(9).uniform(-36, 36, (1500, 3)).astype(np.float32)
    d, i = W.knn(pts, q, k)
    exp = O.knn_brute(1, pts, q, k, threads=8)
    assert np.array_equal(d.view(np.uint32), exp[:, :, 0])
    assert np.array_equal(i, exp[:, :, 1].view(np.int32))


def test_oracle_points_radius_match_ckdtree():
    pts = _cloud(1 << 15, 10)
    q = np.random.default_rng(11).uniform(-36, 36, (1500, 3)).astype(np.float32)
    offs, d, i = W.radius(pts, q, 1.0)
    eo, ek = O.radius_brute(1, pts, q, 1.0, threads=8)
    assert offs[-1] > 0
    assert np.array_equal(offs, eo.astype(np.int64))
    assert np.array_equal(d.view(np.uint32), (ek >> np.uint64(32)).astype(np.uint32))
    assert np.array_equal(i, (ek & np.uint64(0xFFFFFFFF)).astype(np.int32))


# ---- GPU: the reference benches write the dumps -----------------------------------------------

def _run(prog: str, *args: str) -> str:
    path = os.path.join(DROPIN, prog)
    if not os.access(path, os.X_OK):
        pytest.skip(f"{prog} not built (make -C oracle dropin needs the reference sources)")
    print(f"[dropin] {prog} {' '.join(args)}", flush=True)
    r = subprocess.run([path, *args], capture_output=True, text=True, timeout=300)
    out = r.stdout + r.stderr
    print(out[-2000:], flush=True)
    assert r.returncode == 0, out[-2000:]
    return out


@pytest.mark.gpu
def test_cloud_bench_dump_vs_ckdtree(tmp_path):
    import quarry

    path = str(tmp_path / "cloud.qrcd")
    _run("gpu_cloud_bench", "--points", "262144", "--queries", "20000", "--query", "knn", "--k", "8",
         "--runs", "1", "--dump", path)
    pts, q = dumps.read_qrcd(path)
    assert pts.shape == (262144, 3) and q.shape == (20000, 3)
    tree = W._tree(pts)
    cloud = quarry.PointCloud(pts)
    for wide in (False, True):
        d, i = cloud.knn(q, 8, order="morton", wide=wide)
        wd, wi = W.knn(pts, q, 8, tree=tree)
        assert np.array_equal(d.view(np.uint32), wd.view(np.uint32))
        assert np.array_equal(i, wi)
        d1, i1 = cloud.closest(q, wide=wide)
        assert np.array_equal(d1.view(np.uint32), wd[:, 0].view(np.uint32)) and np.array_equal(i1, wi[:, 0])
    r = 0.01 * float(np.linalg.norm(pts.max(axis=0) - pts.min(axis=0)))
    offs, d, i = cloud.radius(q, r)
    wo, wd, wi = W.radius(pts, q, r, tree=tree)
    assert wo[-1] > 0
    assert np.array_equal(offs, wo) and np.array_equal(d.view(np.uint32), wd.view(np.uint32))
    assert np.array_equal(i, wi)


@pytest.mark.gpu
def test_query_bench_dump_vs_oracle(tmp_path):
    import quarry

    path = str(tmp_path / "scene.qryd")
    _run("gpu_query_bench", "--grid", "64", "--points", "5000", "--dump", path)
    tris, q = dumps.read_qryd(path)
    assert tris.shape == (2 * 64 * 64, 3, 3) and q.shape == (5000, 3)
    scene = quarry.Scene(tris)
    items = tris.reshape(-1, 9)
    d, i = scene.knn(q, 16, order="morton", wide=True)
    exp = O.knn_brute(0, items, q, 16, threads=8)
    assert np.array_equal(d.view(np.uint32), exp[:, :, 0]) and np.array_equal(i, exp[:, :, 1].view(np.int32))
    offs, d, i = scene.radius(q, 2.0)
    eo, ek = O.radius_brute(0, items, q, 2.0, threads=8)
    assert np.array_equal(offs, eo.astype(np.int64))
    assert np.array_equal(i, (ek & np.uint64(0xFFFFFFFF)).astype(np.int32))
