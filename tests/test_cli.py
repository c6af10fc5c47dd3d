"""The command line `python -m quarry selftest | run | query` (src/cli/main.cpp).

CPU: its fingerprint functions equal the oracle's restatement of the reference's
FNV-1a fingerprints (src/core/fingerprint.h:16-95) on random records, and the
option surface parses.  GPU: selftest passes; `query` and `run` print the
fingerprint of the oracle's brute force on the same inputs; on a QRYD dump written
by the reference's own gpu_query_bench, `query` prints the hash that bench's CPU
reference prints ("fingerprint cpu ...").
"""
from __future__ import annotations

import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "quarry-gpu_amd")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "dropin")


def _hash(line: str) -> int:
    return int(re.search(r"hash=0x([0-9a-f]{16})", line).group(1), 16)


def test_fingerprints_equal_oracle_hashes():
    from quarry import cli

    rng = np.random.default_rng(5)
    n, k = 300, 6
    d2 = rng.uniform(0, 50, (n, k)).astype(np.float32)
    idx = rng.integers(0, 10**6, (n, k)).astype(np.int32)
    idx[::7, 0] = -1  # not found
    rec = np.stack([d2.view(np.uint32), idx.view(np.uint32)], axis=-1)
    rec[idx < 0, 0] = 0
    h, _ = O.closest_hash(rec[:, :1, :].copy())
    assert _hash(cli.closest_fingerprint(np.where(idx[:, 0] < 0, 0, d2[:, 0]).astype(np.float32),
                                         idx[:, 0])) == h
    full = rng.integers(0, 10**6, (n, k)).astype(np.int32)
    rk = np.stack([d2.view(np.uint32), full.view(np.uint32)], axis=-1)
    h, _ = O.knn_hash(rk, k)
    assert _hash(cli.knn_fingerprint(d2, full)) == h
    counts = rng.integers(0, 5, n)
    offs = np.zeros(n + 1, np.int32)
    np.cumsum(counts, out=offs[1:])
    m = int(offs[-1])
    cd = rng.uniform(0, 4, m).astype(np.float32)
    ci = rng.integers(0, 10**5, m).astype(np.int32)
    keys = (cd.view(np.uint32).astype(np.uint64) << np.uint64(32)) | ci.astype(np.uint32).astype(np.uint64)
    h, _ = O.csr_hash(offs, keys)
    assert _hash(cli.list_fingerprint(offs, cd, ci)) == h
    hit = rng.random(n) < 0.7
    t = rng.uniform(0, 9, n).astype(np.float32)
    res = (hit.astype(np.uint64) << np.uint64(32)) | np.where(hit, t.view(np.uint32), 0).astype(np.uint64)
    h, hits, _ = O.ray_hash(res)
    fp = cli.ray_fingerprint(hit, t)
    assert _hash(fp) == h and f"hits={hits} " in fp


def test_parser_surface():
    from quarry import cli

    a = cli.parser().parse_args(["query", "--type", "knn", "--k", "8", "--scene", "mixed", "--wide",
                                 "--builder", "sah", "--order", "morton", "--stats"])
    assert (a.command, a.type, a.k, a.scene, a.wide, a.builder) == ("query", "knn", 8, "mixed", True, "sah")
    a = cli.parser().parse_args(["run"])
    assert (a.scene, a.grid, a.mode, a.rays, a.runs, a.seed) == ("terrain", 256, "down", 200000, 5, 1)
    with pytest.raises(SystemExit):
        cli.parser().parse_args(["bogus"])


def _cli(*args: str) -> str:
    print("[cli]", " ".join(args), flush=True)
    env = dict(os.environ, PYTHONPATH=os.pathsep.join([PKG, ROOT, os.environ.get("PYTHONPATH", "")]))
    r = subprocess.run([sys.executable, "-m", "quarry", *args], capture_output=True, text=True, timeout=300,
                       env=env, cwd=ROOT)
    out = r.stdout + r.stderr
    print(out[-3000:], flush=True)
    assert r.returncode == 0, out[-2000:]
    return out


@pytest.mark.gpu
def test_selftest_passes():
    out = _cli("selftest")
    assert "SELFTEST PASS (0 failures)" in out and "[FAIL]" not in out


@pytest.mark.gpu
@pytest.mark.parametrize("qtype,extra", [("closest", ()), ("knn", ("--k", "8", "--wide")),
                                         ("radius", ("--radius", "3.0", "--order", "morton"))])
def test_query_fingerprint_equals_oracle(qtype, extra):
    from quarry import cli
    from quarry.demo import _terrain

    out = _cli("query", "--grid", "48", "--points", "2500", "--runs", "1", "--type", qtype, "--stats", *extra)
    tris = _terrain(48)
    pts = cli.make_points(tris, 2500, 1)
    items = tris.reshape(-1, 9)
    if qtype == "closest":
        want, _ = O.closest_hash(O.knn_brute(0, items, pts, 0, threads=8))
    elif qtype == "knn":
        want, _ = O.knn_hash(O.knn_brute(0, items, pts, 8, threads=8), 8)
    else:
        want, _ = O.csr_hash(*O.radius_brute(0, items, pts, 3.0, threads=8))
    line = [ln for ln in out.splitlines() if ln.startswith("fingerprint ")][0]
    assert _hash(line) == want
    assert "traversal avg_node_tests=" in out


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["down", "random"])
def test_run_fingerprint_equals_oracle(mode):
    from quarry import cli
    from quarry.demo import _terrain

    out = _cli("run", "--grid", "48", "--rays", "4000", "--runs", "2", "--mode", mode, "--wide")
    tris = _terrain(48)
    o, d, tmax = cli.make_rays(tris, mode, 4000, 1)
    rays = np.zeros((4000, 8), np.float32)
    rays[:, 0:3], rays[:, 3:6], rays[:, 6], rays[:, 7] = o, d, np.float32(1e-6), np.float32(tmax)
    want, hits, _ = O.ray_hash(O.ray_brute(tris.reshape(-1, 9), rays, threads=8))
    line = [ln for ln in out.splitlines() if ln.startswith("fingerprint ")][0]
    assert _hash(line) == want and f"hits={hits} " in line


@pytest.mark.gpu
@pytest.mark.parametrize("qtype,extra", [("knn", ("--k", "16")), ("radius", ("--radius", "2.0"))])
def test_query_on_reference_dump(tmp_path, qtype, extra):
    prog = os.path.join(DROPIN, "gpu_query_bench")
    if not os.access(prog, os.X_OK):
        pytest.skip("gpu_query_bench not built (make -C oracle dropin needs the reference sources)")
    dump = str(tmp_path / "scene.qryd")
    common = ["--grid", "64", "--points", "3000", "--query", qtype, *extra]
    subprocess.run([prog, *common, "--dump", dump], check=True, capture_output=True, timeout=120)
    r = subprocess.run([prog, *common, "--runs", "1", "--warmup-ms", "0", "--verify"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    cpu = [ln for ln in r.stdout.splitlines() if ln.startswith("fingerprint cpu ")][0]
    out = _cli("query", "--input", dump, "--runs", "1", "--type", qtype, *extra)
    line = [ln for ln in out.splitlines() if ln.startswith("fingerprint ")][0]
    assert _hash(line) == _hash(cpu)
