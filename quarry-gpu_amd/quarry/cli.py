"""Command line of the GPU path: `python -m quarry selftest | run | query`.

Mirrors the reference's CLI (src/cli/main.cpp:30-332): the same commands, options
and output lines -- scene and timing lines, then a `fingerprint ...` line in the
reference's format (src/core/fingerprint.h:16-95: FNV-1a over per-ray (hit << 32 |
t bits), or per query over (found << 32 | dist2 bits, index + 1) for closest and
(length, then dist2 bits and index + 1 per entry) for lists).  The engine is the
GPU (binary or 4-wide nodes, LBVH or SAH); `--input FILE` reads a QRYD dump
(bench/gpu_query_bench.cpp:232-259), so a fingerprint can be set beside the one
the reference prints for the same bytes.

selftest runs checks that need no CPU reference: closed-form answers on tiny
scenes, binary == wide == Morton-ordered results, LBVH == SAH results, repeatable
tree hashes, validate() on every scene type.
"""
from __future__ import annotations

import argparse
import math
import sys
import time

import numpy as np

from . import _abi as A
from .api import _ORDERS, _RAY_ORDERS, GaussianCloud, PointCloud, Scene, _order
from .demo import F, Pcg32, _city, _random_tris, _terrain
from .dumps import read_qryd

FNV_OFFSET = 0xCBF29CE484222325
FNV_PRIME = 0x100000001B3
MASK64 = (1 << 64) - 1


# ---- fingerprints (src/core/fingerprint.h) ---------------------------------------------------

def ray_fingerprint(hit: np.ndarray, t: np.ndarray) -> str:
    bits = np.where(hit, np.ascontiguousarray(t, np.float32).view(np.uint32), 0).astype(np.uint64)
    words = (hit.astype(np.uint64) << np.uint64(32)) | bits
    h = FNV_OFFSET
    for v in words.tolist():
        h = ((h ^ v) * FNV_PRIME) & MASK64
    sum_t = float(np.sum(t[hit].astype(np.float64)))
    return f"rays={hit.size} hits={int(hit.sum())} sum_t={sum_t:.6f} hash=0x{h:016x}"


def closest_fingerprint(d2: np.ndarray, idx: np.ndarray) -> str:
    found = idx >= 0
    bits = np.where(found, np.ascontiguousarray(d2, np.float32).view(np.uint32), 0).astype(np.uint64)
    a = (found.astype(np.uint64) << np.uint64(32)) | bits
    b = np.where(found, idx.astype(np.int64) + 1, 0).astype(np.uint64)
    h = FNV_OFFSET
    for x, y in zip(a.tolist(), b.tolist()):
        h = ((h ^ x) * FNV_PRIME) & MASK64
        h = ((h ^ y) * FNV_PRIME) & MASK64
    s = float(np.sum(d2[found].astype(np.float64)))
    return f"queries={idx.size} results={int(found.sum())} sum_dist2={s:.6f} hash=0x{h:016x}"


def list_fingerprint(offsets: np.ndarray, d2: np.ndarray, idx: np.ndarray) -> str:
    bits = np.ascontiguousarray(d2, np.float32).view(np.uint32).tolist()
    ids = (idx.astype(np.int64) + 1).tolist()
    h = FNV_OFFSET
    offs = offsets.tolist()
    for q in range(len(offs) - 1):
        lo, hi = offs[q], offs[q + 1]
        h = ((h ^ (hi - lo)) * FNV_PRIME) & MASK64
        for j in range(lo, hi):
            h = ((h ^ bits[j]) * FNV_PRIME) & MASK64
            h = ((h ^ ids[j]) * FNV_PRIME) & MASK64
    s = float(np.sum(d2.astype(np.float64)))
    return f"queries={len(offs) - 1} results={len(ids)} sum_dist2={s:.6f} hash=0x{h:016x}"


def knn_fingerprint(d2: np.ndarray, idx: np.ndarray) -> str:
    """kNN rows hold at most k entries; missing ones (fewer than k items) are -1."""
    valid = idx >= 0
    counts = valid.sum(axis=1)
    offsets = np.zeros(idx.shape[0] + 1, np.int64)
    np.cumsum(counts, out=offsets[1:])
    return list_fingerprint(offsets, d2[valid], idx[valid])


# ---- inputs (src/core/scene.cpp:128-209) -----------------------------------------------------

def build_scene(args) -> np.ndarray:
    if args.input:
        tris, _ = read_qryd(args.input)
        return tris
    if args.scene == "terrain":
        return _terrain(args.grid)
    if args.scene == "random":
        return _random_tris(args.random_tris, args.seed)
    if args.scene == "mixed":
        return np.concatenate([_terrain(args.grid), _random_tris(args.random_tris, args.seed)])
    if args.scene == "city":
        return _city(args.buildings, args.seed)
    raise ValueError(f"unknown scene {args.scene}")


def _bounds(tris: np.ndarray):
    v = tris.reshape(-1, 3)
    return v.min(axis=0).astype(np.float32), v.max(axis=0).astype(np.float32)


def make_rays(tris: np.ndarray, mode: str, count: int, seed: int):
    """(origins, dirs, tmax) as the reference draws them (scene.cpp:146-189)."""
    mn, mx = _bounds(tris)
    rng = Pcg32(seed, 2)
    center = (mn + mx) * F(0.5)
    ext = mx - mn
    o = np.zeros((count, 3), np.float32)
    d = np.zeros((count, 3), np.float32)
    if mode == "down":
        for i in range(count):
            o[i] = [rng.uniform(mn[0], mx[0]), mx[1] + F(10.0), rng.uniform(mn[2], mx[2])]
        d[:, 1] = -1.0
        return o, d, float(ext[1] + F(20.0))
    radius = F(1.6) * F(np.sqrt(np.dot(ext, ext)))
    spread = F(0.9)
    for i in range(count):
        theta = rng.uniform(0.0, 6.2831853)
        phi = F(math.acos(float(rng.uniform(-1.0, 1.0))))
        sp = F(math.sin(phi))
        o[i] = [center[0] + radius * sp * F(math.cos(theta)), center[1] + radius * sp * F(math.sin(theta)),
                center[2] + radius * F(math.cos(phi))]
        target = np.array([center[j] + rng.uniform(-1.0, 1.0) * ext[j] * spread for j in range(3)], np.float32)
        v = target - o[i]
        d[i] = v / F(np.sqrt(np.dot(v, v)))
    return o, d, float(radius * F(4.0))


def make_points(tris: np.ndarray, count: int, seed: int) -> np.ndarray:
    """Uniform in the scene box grown to 1.2x (scene.cpp:195-209)."""
    mn, mx = _bounds(tris)
    rng = Pcg32(seed, 3)
    center = (mn + mx) * F(0.5)
    half = (mx - mn) * F(0.6)
    out = np.zeros((count, 3), np.float32)
    for i in range(count):
        out[i] = [center[j] + rng.uniform(-half[j], half[j]) for j in range(3)]
    return out


# ---- commands --------------------------------------------------------------------------------

def _time_runs(fn, runs: int):
    times, out = [], None
    for r in range(max(1, runs)):
        t0 = time.perf_counter()
        res = fn()
        times.append(1000.0 * (time.perf_counter() - t0))
        if r == 0:
            out = res
    times.sort()
    return out, times


def _header(args, tris, scene_ms, scene: Scene, what: str) -> None:
    print(f"scene={'file' if args.input else args.scene} tris={tris.shape[0]} {what} seed={args.seed} "
          f"engine=gpu builder={args.builder} nodes={'wide4' if args.wide else 'binary'}")
    print(f"scene_ms={scene_ms:.2f} build_ms={scene.build_ms:.2f} nodes={scene.node_count} depth={scene.depth}")


def cmd_run(args) -> int:
    t0 = time.perf_counter()
    tris = build_scene(args)
    o, d, tmax = make_rays(tris, args.mode, args.rays, args.seed)
    scene_ms = 1000.0 * (time.perf_counter() - t0)
    scene = Scene(tris, builder=args.builder)
    if args.wide:
        scene.build_wide()
    (hit, t), times = _time_runs(lambda: scene.raycast(o, d, tmax=tmax, order=args.order, wide=args.wide),
                                 args.runs)
    med = times[len(times) // 2]
    _header(args, tris, scene_ms, scene, f"rays={args.rays} mode={args.mode}")
    print(f"runs={len(times)} median_ms={med:.2f} min_ms={times[0]:.2f} max_ms={times[-1]:.2f} "
          f"Mrays_s={args.rays / (med * 1e3) if med > 0 else 0.0:.2f}")
    print(f"fingerprint {ray_fingerprint(hit, t)}")
    if args.stats:
        rays = np.zeros((o.shape[0], 8), np.float32)
        rays[:, 0:3], rays[:, 3:6], rays[:, 6], rays[:, 7] = o, d, np.float32(1e-6), np.float32(tmax)
        _, _, st = scene._bvh.raycast(rays, _order(args.order, _RAY_ORDERS), args.wide, stats=True)
        _print_stats(st, o.shape[0])
    return 0


def _print_stats(st, n: int) -> None:
    n = max(1, n)
    print(f"traversal avg_node_tests={st.node_tests / n:.1f} avg_tri_tests={st.leaf_tests / n:.1f} "
          f"avg_nodes_opened={st.nodes_opened / n:.1f}")


def cmd_query(args) -> int:
    t0 = time.perf_counter()
    tris = build_scene(args)
    if args.input:
        _, pts = read_qryd(args.input)
        if args.points >= 0:
            pts = pts[: args.points]
    else:
        pts = make_points(tris, 20000 if args.points < 0 else args.points, args.seed)
    scene_ms = 1000.0 * (time.perf_counter() - t0)
    scene = Scene(tris, builder=args.builder)
    if args.wide:
        scene.build_wide()
    if args.type == "closest":
        fn = lambda: scene.closest(pts, order=args.order, wide=args.wide)  # noqa: E731
    elif args.type == "knn":
        fn = lambda: scene.knn(pts, args.k, order=args.order, wide=args.wide)  # noqa: E731
    else:
        fn = lambda: scene.radius(pts, args.radius, order=args.order, wide=args.wide)  # noqa: E731
    res, times = _time_runs(fn, args.runs)
    med = times[len(times) // 2]
    extra = {"knn": f" k={args.k}", "radius": f" radius={args.radius:.4f}"}.get(args.type, "")
    _header(args, tris, scene_ms, scene, f"points={pts.shape[0]} type={args.type}{extra}")
    if args.type == "closest":
        fp = closest_fingerprint(*res)
        nres = int((res[1] >= 0).sum())
    elif args.type == "knn":
        fp = knn_fingerprint(*res)
        nres = int((res[1] >= 0).sum())
    else:
        fp = list_fingerprint(*res)
        nres = int(res[0][-1])
    print(f"runs={len(times)} median_ms={med:.2f} min_ms={times[0]:.2f} max_ms={times[-1]:.2f} "
          f"Mqueries_s={pts.shape[0] / (med * 1e3) if med > 0 else 0.0:.2f} "
          f"results_per_query={nres / max(1, pts.shape[0]):.2f}")
    print(f"fingerprint {fp}")
    if args.stats:
        b, o = scene._bvh, _order(args.order, _ORDERS)
        if args.type == "closest":
            st = b.closest(pts, o, args.wide, stats=True)[2]
        elif args.type == "knn":
            st = b.knn(pts, args.k, o, args.wide, stats=True)[2]
        else:
            st = b.radius(pts, args.radius, o, args.wide, stats=True)[3]
        _print_stats(st, pts.shape[0])
    return 0


class _Checks:
    def __init__(self):
        self.failures = 0

    def __call__(self, name: str, ok: bool, detail: str = "") -> None:
        print(f"[{'PASS' if ok else 'FAIL'}] {name}{'' if ok else ' ' + detail}", flush=True)
        if not ok:
            self.failures += 1


def _same(a, b) -> bool:
    return all(np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8)) for x, y in zip(a, b))


def cmd_selftest(args) -> int:
    """Closed-form answers and self-consistency (src/cpu/selftest.cpp:534-550 covers
    the same ground on the CPU)."""
    check = _Checks()
    # known answers: the triangle z = 0 over (0,0) (4,0) (0,4), and its mirror
    tri = np.array([[[0, 0, 0], [4, 0, 0], [0, 4, 0]], [[10, 0, 0], [14, 0, 0], [10, 4, 0]]], np.float32)
    s = Scene(tri)
    check("validate tiny scene", s.validate() == "", s.validate())
    hit, t = s.raycast(np.array([[1, 1, 5], [11, 1, -2], [5, 5, 5]], np.float32),
                       np.array([[0, 0, -1], [0, 0, 1], [0, 0, -1]], np.float32))
    check("ray known answers", hit.tolist() == [True, True, False] and t[0] == 5.0 and t[1] == 2.0
          and math.isinf(t[2]), f"{hit} {t}")
    q = np.array([[1, 1, 3], [-3, 0, 0], [2, 2, 0], [12, 1, -1]], np.float32)
    d2, idx = s.closest(q)
    check("closest known answers", d2.tolist() == [9.0, 9.0, 0.0, 1.0] and idx.tolist() == [0, 0, 0, 1],
          f"{d2} {idx}")
    kd, kid = s.knn(q[:1], 2)
    # (1,1,3) to the second triangle: nearest point (10,1,0), squared distance 81 + 9
    check("knn known answers", kd[0].tolist() == [9.0, 90.0] and kid[0].tolist() == [0, 1], f"{kd} {kid}")
    offs, rd2, rid = s.radius(q, 3.0)
    check("radius known answers", offs.tolist() == [0, 1, 2, 3, 4] and rid.tolist() == [0, 0, 0, 1],
          f"{offs} {rid}")
    pc = PointCloud(np.array([[0, 0, 0], [3, 4, 0], [0, 0, 2]], np.float32))
    pd2, pidx = pc.knn(np.array([[0, 0, 0]], np.float32), 3)
    check("point knn known answers", pd2[0].tolist() == [0.0, 4.0, 25.0] and pidx[0].tolist() == [0, 2, 1])
    gc = GaussianCloud(np.zeros((2, 3), np.float32) + np.array([[0, 0, 0], [10, 0, 0]], np.float32),
                       np.array([[0, 0, 0, 1], [0, 0, 0, 1]], np.float32),
                       np.array([[1, 2, 3], [1, 1, 1]], np.float32))
    gd2, gidx = gc.closest(np.array([[0, 0, 5], [0, 0, 1], [12, 0, 0]], np.float32))
    # on an axis the distance is the gap to the semi-axis: 5 - 3 and 12 - 11 (the root
    # solve may leave the last place)
    check("gaussian known answers", np.allclose(gd2, [4.0, 0.0, 1.0], rtol=1e-6, atol=0)
          and gidx.tolist() == [0, 0, 1], f"{gd2} {gidx}")

    # self-consistency on a terrain, a city, a mixed scene
    for name, tris in (("terrain", _terrain(48)), ("city", _city(400, 3)),
                       ("mixed", np.concatenate([_terrain(24), _random_tris(500, 5)]))):
        lb = Scene(tris)
        sah = Scene(tris, builder="sah")
        check(f"{name} validate", lb.validate() == "" and sah.validate() == "")
        pts = make_points(tris, 3000, 11)
        o, d, tmax = make_rays(tris, "random", 3000, 12)
        base = [lb.closest(pts), lb.knn(pts, 8), lb.radius(pts, 2.0), lb.raycast(o, d, tmax=tmax)]
        for label, sc, wide, order in (("wide", lb, True, "morton"), ("sah", sah, False, "input"),
                                       ("sah wide", sah, True, "morton")):
            got = [sc.closest(pts, order=order, wide=wide), sc.knn(pts, 8, order=order, wide=wide),
                   sc.radius(pts, 2.0, order=order, wide=wide), sc.raycast(o, d, tmax=tmax, wide=wide)]
            check(f"{name} {label} == binary lbvh", all(_same(a, b) for a, b in zip(base, got)))
        h1 = A.Bvh(A.PRIM_TRI, tris.reshape(-1, 9), 0)
        h2 = A.Bvh(A.PRIM_TRI, tris.reshape(-1, 9), 0)
        check(f"{name} tree hash repeatable", h1.tree_hash() == h2.tree_hash())
        h1.close()
        h2.close()
    print(f"\n{'SELFTEST PASS' if check.failures == 0 else 'SELFTEST FAIL'} "
          f"({check.failures} failure{'' if check.failures == 1 else 's'})")
    return 0 if check.failures == 0 else 3


def parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(prog="python -m quarry", description=__doc__.split("\n\n")[0])
    p.add_argument("command", choices=["selftest", "run", "query"])
    p.add_argument("--scene", default="terrain", choices=["terrain", "random", "mixed", "city"])
    p.add_argument("--grid", type=int, default=256, help="terrain cells per side (256: 131,072 triangles)")
    p.add_argument("--random-tris", type=int, default=0)
    p.add_argument("--buildings", type=int, default=65536)
    p.add_argument("--input", default="", help="QRYD dump: its triangles (and points for query)")
    p.add_argument("--builder", default="lbvh", choices=["lbvh", "sah"])
    p.add_argument("--wide", action="store_true", help="traverse 4-wide nodes")
    p.add_argument("--order", default="input")
    p.add_argument("--runs", type=int, default=5)
    p.add_argument("--seed", type=int, default=1)
    p.add_argument("--stats", action="store_true")
    p.add_argument("--rays", type=int, default=200000)
    p.add_argument("--mode", default="down", choices=["down", "random"])
    p.add_argument("--type", default="closest", choices=["closest", "knn", "radius"])
    p.add_argument("--points", type=int, default=-1)
    p.add_argument("--k", type=int, default=16)
    p.add_argument("--radius", type=float, default=2.0)
    return p


def main(argv=None) -> int:
    args = parser().parse_args(argv)
    if args.scene == "random" and args.random_tris == 0:
        args.random_tris = 200000
    if args.scene == "mixed" and args.random_tris == 0:
        args.random_tris = 50000
    if args.type == "knn" and not 1 <= args.k <= 32:
        print("k must be in [1, 32]", file=sys.stderr)
        return 1
    return {"selftest": cmd_selftest, "run": cmd_run, "query": cmd_query}[args.command](args)


if __name__ == "__main__":
    sys.exit(main())
