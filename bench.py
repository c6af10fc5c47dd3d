#!/usr/bin/env python3
"""Benchmark of the BVH query path on one MI355X.

Workload (the configuration the headline metric is quoted on): a 1024 x 1024-cell
z-up terrain (2,097,152 FP32 triangles) with 100,000 query points uniform in the
scene box grown by 4 on every side.  One step is one serving pass over the
device-resident index (built with binned SAH, the reference's builder="sah";
--builder lbvh serves from the LBVH tree): closest point, k-nearest (k = 16) and
radius (r = 2.0) for all 100,000 queries, Morton-ordered, on 4-wide nodes.  value =
queries answered per second over the K timed steps (300,000 per step), in millions.
The same pass on the other tree is reported beside it (same result bits).

Around the timed region the same run also measures and checks:
  * the LBVH build of the 2M-triangle scene (median of several builds) with its
    HBM roofline, the 4-wide collapse, and the SAH build;
  * a dirty refit after moving 1% of the grid vertices by N(0, 0.05);
  * closest-hit rays: 1,000,000 down rays on the 256 x 256-cell terrain;
  * bit-exact parity of every result set against the CPU oracle;
  * the reference CPU path (or the oracle restatement when the reference
    build is absent) timed on a bounded sample of the same queries.
Progress goes to stderr; stdout carries one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "quarry-gpu_amd"))

HBM_PEAK_GBS = 8000.0
TRI_BYTES = 36


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--queries", type=int, default=100_000)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--radius", type=float, default=2.0)
    ap.add_argument("--builder", choices=["sah", "lbvh"], default="sah",
                    help="builder of the index the timed serving pass runs on")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity legs")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--profile", action="store_true", help="short run for rocprof: no CPU legs, no extras")
    ap.add_argument("--no-extras", action="store_true", help="skip the C3 / C4 / C5 legs")
    return ap.parse_args()


def main() -> int:
    import faulthandler

    faulthandler.enable()
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus above 1 is not "
                                   "supported (the path is measured on one GPU's share)",
                          "n_gpus": args.gpus}))
        return 2

    from quarry import _abi as A
    from quarry import scenes as S

    dev = A.describe()
    if dev is None:
        log("no HIP device visible")
        return 1
    log("device:", dev)
    A.check(A.lib().qr_warmup())
    if args.profile:
        calibration_copy(A)

    # ---- workload ---------------------------------------------------------------
    t0 = time.time()
    tris, queries = S.workload_c2(seed=1, cells=args.cells, queries=args.queries)
    n = tris.shape[0]
    log(f"scene: {n} triangles, {queries.shape[0]} queries ({time.time() - t0:.1f}s to generate)")

    dtris = A.DeviceArray(tris)
    builds, wides = [], []
    bvh = None
    # The timed builds record no phase events (each event between two kernels idles
    # the device for ~12 us); one more build with them gives the phase split.
    A.check(A.lib().qr_set_build_phase_timing(0))
    try:
        for i in range(5 if not args.profile else 2):
            if bvh is not None:
                bvh.close()
            bvh = A.Bvh.from_device(A.PRIM_TRI, dtris.ptr.value, n)
            builds.append(bvh.timings)
            # the collapse of each build: the first one in the process also loads its
            # kernels, so the median is the steady-state figure, as for the build
            wides.append(bvh.build_wide())
    finally:
        A.check(A.lib().qr_set_build_phase_timing(1))
    build_ms = statistics.median(t.total_ms for t in builds)
    probe = A.Bvh.from_device(A.PRIM_TRI, dtris.ptr.value, n)
    bt = probe.timings
    probe.close()
    wide_ms = statistics.median(wides)
    info = bvh.info()
    log(f"build: {build_ms:.3f} ms (with phase events {bt.total_ms:.3f}: morton {bt.morton_ms:.3f}, sort {bt.sort_ms:.3f}, "
        f"tree {bt.tree_ms:.3f}, leaves+boxes {bt.refit_ms:.3f}); depth {info.depth}; "
        f"wide collapse {wide_ms:.3f} ms ({info.wide_count} nodes)")

    # The index the serving pass runs on: binned SAH by default (the reference's
    # builder="sah"; the results are the same bits on either tree, and on this terrain
    # the SAH tree opens about a third of the LBVH tree's nodes per query).  Both
    # trees stay resident; the LBVH tree is the one the build roofline and the refit
    # are measured on, and its serving pass is reported beside the headline.
    lbvh = bvh
    sah = A.Bvh.from_device(A.PRIM_TRI, dtris.ptr.value, n, A.BUILDER_SAH)
    # the first SAH build also sizes its workspaces; the median of the next five is
    # the figure (as for the LBVH build)
    sah_builds = []
    A.check(A.lib().qr_set_build_phase_timing(0))
    try:
        for _ in range(5 if not args.profile else 1):
            sah.close()
            sah = A.Bvh.from_device(A.PRIM_TRI, dtris.ptr.value, n, A.BUILDER_SAH)
            sah_builds.append(sah.timings)
    finally:
        A.check(A.lib().qr_set_build_phase_timing(1))
    sah_build_ms = statistics.median(t.total_ms for t in sah_builds)
    sah_work = A.last_sah_stats()
    sah_wide_ms = sah.build_wide()
    # the phase split from one build with the phase events, and the split of the level
    # loop from one more with the per-level events too (those cost about 0.2 ms)
    probe = A.Bvh.from_device(A.PRIM_TRI, dtris.ptr.value, n, A.BUILDER_SAH)
    sst = probe.timings
    probe.close()
    sah_levels_ms = sst.tree_ms
    sah_split_total_ms = sst.total_ms
    A.check(A.lib().qr_set_build_phase_timing(2))
    try:
        probe = A.Bvh.from_device(A.PRIM_TRI, dtris.ptr.value, n, A.BUILDER_SAH)
        pt = probe.timings
        probe.close()
    finally:
        A.check(A.lib().qr_set_build_phase_timing(1))
    # levels: the large tasks' level loop (binning, plane choice, partition + child
    # routing) and the local subtrees + serial tails (scatter_ms)
    sah_parts = {"prep_ms": round(sst.morton_ms, 4), "levels_ms": round(sst.tree_ms, 4),
                 "subtrees_ms": round(sst.scatter_ms, 4), "leaves_boxes_ms": round(sst.refit_ms, 4),
                 "with_level_events": {"total_ms": round(pt.total_ms, 4), "bin_ms": round(pt.sah_bin_ms, 4),
                                       "split_ms": round(pt.sah_split_ms, 4),
                                       "partition_ms": round(pt.sah_part_ms, 4)}}
    log(f"SAH build: {sah_build_ms:.3f} ms (with phase events {sst.total_ms:.3f}: prep {sst.morton_ms:.3f}, "
        f"levels {sst.tree_ms:.3f} incl. local subtrees {sst.scatter_ms:.3f}; leaves+boxes {sst.refit_ms:.3f}); depth {sah.info().depth}; "
        f"wide collapse {sah_wide_ms:.3f} ms; with per-level events {pt.total_ms:.3f} ms: large tasks bin "
        f"{pt.sah_bin_ms:.3f}, split {pt.sah_split_ms:.3f}, partition {pt.sah_part_ms:.3f}")
    serve = sah if args.builder == "sah" else lbvh

    pb1 = A.PointBuffer(queries, 1)
    pbk = A.PointBuffer(queries, args.k)

    def step(tree):
        a = pb1.closest(tree)
        sa = A.last_query_stats().kernel_ms
        b = pbk.knn(tree, args.k)
        sb = A.last_query_stats().kernel_ms
        c = pb1.radius(tree, args.radius)
        sc = A.last_query_stats().kernel_ms
        return (a, b, c), (sa, sb, sc)

    def timed(tree, steps, warmup):
        for _ in range(warmup):
            step(tree)
        A.check(A.lib().qr_dev_synchronize())
        call_ms, kern_ms = [], []
        t_start = time.perf_counter()
        for _ in range(steps):
            c, k = step(tree)
            call_ms.append(c)
            kern_ms.append(k)
        A.check(A.lib().qr_dev_synchronize())
        wall = time.perf_counter() - t_start
        names = ("closest", "knn", "radius")
        return (wall, {name: statistics.median(c[i] for c in call_ms) for i, name in enumerate(names)},
                {name: statistics.mean(k[i] for k in kern_ms) for i, name in enumerate(names)})

    q_per_step = 3 * queries.shape[0]
    wall, per_op, per_kernel = timed(serve, args.steps, args.warmup)
    ms_per_step = 1000.0 * wall / args.steps
    value = q_per_step / (wall / args.steps) / 1e6
    log(f"step ({args.builder} index): {ms_per_step:.3f} ms wall; per call (event ms) {per_op}; "
        f"traversal kernels {per_kernel}")
    other = lbvh if serve is sah else sah
    other_index = None
    if not args.profile:  # (a profile covers the serving pass only)
        o_steps = max(3, args.steps // 2)
        o_wall, o_per_op, o_per_kernel = timed(other, o_steps, 2)
        other_index = {"builder": "lbvh" if other is lbvh else "sah",
                       "mq_per_s": round(q_per_step / (o_wall / o_steps) / 1e6, 3),
                       "ms_per_step": round(1000.0 * o_wall / o_steps, 4),
                       "per_op_kernel_ms": {k: round(v, 4) for k, v in o_per_kernel.items()}}
        log(f"other index: {other_index}")
    bvh = serve

    # traversal work per query (one extra counted pass, outside the timed region)
    work = {}
    nq = queries.shape[0]
    for name, fn in (("closest", lambda: pb1.closest(bvh, stats=True)),
                     ("knn", lambda: pbk.knn(bvh, args.k, stats=True)),
                     ("radius", lambda: pb1.radius(bvh, args.radius, stats=True))):
        fn()
        st = A.last_query_stats()
        work[name] = (st.node_tests / nq, st.leaf_tests / nq, st.nodes_opened / nq)
    log("node/leaf tests, nodes opened per query: " +
        ", ".join(f"{k} {v[0]:.1f}/{v[1]:.1f}/{v[2]:.1f}" for k, v in work.items()))
    # The same counts for a sequential nearest-first walk (the four-lane kernels,
    # traversal mode 2): the work the algorithm needs, which the roofline counts.
    # The serving kernels may open more nodes than that (a wave expands 16 of one
    # query's nodes per round); their extra reads are not counted as achieved bytes.
    seq = {}
    A.check(A.lib().qr_set_traversal_mode(2))
    try:
        for name, fn in (("closest", lambda: pb1.closest(bvh, stats=True)),
                         ("knn", lambda: pbk.knn(bvh, args.k, stats=True)),
                         ("radius", lambda: pb1.radius(bvh, args.radius, stats=True))):
            fn()
            st = A.last_query_stats()
            seq[name] = (st.node_tests / nq, st.leaf_tests / nq, st.nodes_opened / nq)
    finally:
        A.check(A.lib().qr_set_traversal_mode(0))
    log("sequential walk: " + ", ".join(f"{k} {v[0]:.1f}/{v[1]:.1f}/{v[2]:.1f}" for k, v in seq.items()))

    # the same three queries on binary nodes (reported beside the wide pass)
    binary = {}
    for name, fn in (("closest", lambda: pb1.closest(bvh, wide=False)),
                     ("knn", lambda: pbk.knn(bvh, args.k, wide=False)),
                     ("radius", lambda: pb1.radius(bvh, args.radius, wide=False))):
        fn()
        ks = []
        for _ in range(5):
            fn()
            ks.append(A.last_query_stats().kernel_ms)
        binary[name] = round(statistics.median(ks), 4)
    log(f"binary-node traversal kernels (ms): {binary}")
    # the binary pass's results, compared with the oracle below
    pb1.closest(bvh, wide=False)
    bin_c = pb1.read()
    pbk.knn(bvh, args.k, wide=False)
    bin_k = pbk.read()
    _, bin_o, bin_keys = pb1.radius(bvh, args.radius, wide=False, keep=True)

    # ---- roofline of the dominant kernel ------------------------------------------
    dom = max(per_kernel, key=per_kernel.get)
    _, offs, keys = pb1.radius(bvh, args.radius, keep=True)
    radius_total = int(offs[-1])
    # Algorithmic bytes per launch: the bytes the traversal has to read -- each
    # opened 4-wide node (128 B), each tested leaf (36 B triangle + 4 B prim id),
    # the query (12 B) and its order slot (4 B) -- plus the results it writes.
    # Node and leaf counts are the sequential walk's (kernel counters above).
    opened, leaves = seq[dom][2] * nq, seq[dom][1] * nq
    touched = opened * 128 + leaves * (TRI_BYTES + 4) + nq * (12 + 4)
    alg = {
        "closest": touched + nq * 8,
        "knn": touched + nq * 8 * args.k,
        # one walk: counts and captured hits written, then copied into the CSR
        # (offsets and keys); queries with more hits than the capture slot walk again
        "radius": touched + nq * 8 + radius_total * 16,
    }[dom]
    achieved = alg / (per_kernel[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": traffic_for(dom, args),
                "alg_bytes_per_launch": int(alg), "avg_kernel_ms": round(per_kernel[dom], 4)}

    # Build roofline: 5S + 232 bytes per primitive (S = 36 for triangles).
    build_alg = n * (5 * TRI_BYTES + 232)
    build_gbs = build_alg / (build_ms * 1e-3) / 1e9
    # SAH split loop: (S + 8) bytes per slot binned and 16 per slot partitioned, summed
    # over the levels the build ran (slot counts from the build's own level tables),
    # over the levels' time.  The implementation moves 40-byte primitive records
    # instead (40 per slot binned, 80 per slot partitioned): record_bytes.
    sah_alg = sah_work.binned_slots * (TRI_BYTES + 8) + sah_work.partitioned_slots * 16
    sah_gbs = sah_alg / (sah_levels_ms * 1e-3) / 1e9
    sah_roof = {"bound": "hbm", "achieved": round(sah_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(sah_gbs / HBM_PEAK_GBS, 4), "alg_bytes": int(sah_alg), "levels": sah_work.levels,
                "binned_slots": int(sah_work.binned_slots), "partitioned_slots": int(sah_work.partitioned_slots),
                "serial_slots": int(sah_work.serial_slots),
                "record_bytes": int(sah_work.binned_slots * 40 + sah_work.partitioned_slots * 80),
                "levels_ms": round(sah_levels_ms, 4)}
    build = {"ms": round(build_ms, 4), "ms_with_phase_events": round(bt.total_ms, 4),
             "morton_ms": round(bt.morton_ms, 4),
             "sort_ms": round(bt.sort_ms, 4), "tree_ms": round(bt.tree_ms, 4),
             "leaves_boxes_ms": round(bt.refit_ms, 4), "wide_ms": round(wide_ms, 4),
             "depth": info.depth, "sah_ms": round(sah_build_ms, 4),
             "sah_ms_with_phase_events": round(sah_split_total_ms, 4), "sah_phases": sah_parts,
             "sah_roofline": sah_roof, "sah_wide_ms": round(sah_wide_ms, 4),
             "roofline": {"bound": "hbm", "achieved": round(build_gbs, 2), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(build_gbs / HBM_PEAK_GBS, 4),
                          "alg_bytes": build_alg}}

    extra = {}
    parity = {}
    cpu_baseline = None
    if args.profile:
        # the refit kernels for the counter runs: the 1% dirty refit (dense path) and
        # a full refit from device-resident items, then back to the original scene
        moved, dirty = S.displace_vertices(args.cells, tris, 0.01, 0.05, seed=4)
        lbvh.refit(moved, dirty)
        lbvh.refit_device(dtris.ptr.value, n)
        log("profile: dirty refit and full device refit done")
    if not args.profile:
        # ---- parity of the query pass against the CPU oracle -----------------------
        import oracle as O

        threads = min(16, os.cpu_count() or 1)
        log("parity: building the CPU oracle BVH")
        t0 = time.time()
        cb = O.CpuBvh(0, tris)
        res_c = pb1.read()
        pb1.closest(bvh)
        res_c = pb1.read()
        pbk.knn(bvh, args.k)
        res_k = pbk.read()
        exp_c = cb.knn(queries, 0, threads)
        exp_k = cb.knn(queries, args.k, threads)
        exp_o, exp_keys = cb.radius(queries, args.radius, threads)
        pb1.closest(other)
        oth_c = pb1.read()
        pbk.knn(other, args.k)
        oth_k = pbk.read()
        _, oth_o, oth_keys = pb1.radius(other, args.radius, keep=True)
        parity["other_index_same_bits"] = bool(np.array_equal(oth_c, res_c) and np.array_equal(oth_k, res_k)
                                               and np.array_equal(oth_o, offs) and np.array_equal(oth_keys, keys))
        parity["closest"] = bool(np.array_equal(res_c, exp_c))
        parity["knn"] = bool(np.array_equal(res_k, exp_k))
        parity["radius"] = bool(np.array_equal(offs, exp_o) and np.array_equal(keys, exp_keys))
        parity["binary"] = bool(np.array_equal(bin_c, exp_c) and np.array_equal(bin_k, exp_k)
                                and np.array_equal(bin_o, exp_o) and np.array_equal(bin_keys, exp_keys))
        fp = {"closest": hex(O.closest_hash(res_c)[0]), "knn": hex(O.knn_hash(res_k, args.k)[0]),
              "radius": hex(O.csr_hash(offs, keys)[0])}
        log(f"parity {parity} ({time.time() - t0:.1f}s); fingerprints {fp}")
        extra["fingerprints"] = fp
        del cb

        # ---- dirty refit after moving 1% of the vertices ---------------------------
        # Closest, kNN and radius on binary and 4-wide nodes of both trees after the
        # dirty refit, against the CPU BVH of the moved scene.
        moved, dirty = S.displace_vertices(args.cells, tris, 0.01, 0.05, seed=4)
        rt = lbvh.refit(moved, dirty)
        why = lbvh.validate(moved)
        rt_sah = sah.refit(moved, dirty)
        why_sah = sah.validate(moved)
        cb2 = O.CpuBvh(0, moved)
        mv_c, mv_k = cb2.knn(queries, 0, threads), cb2.knn(queries, args.k, threads)
        mv_o, mv_keys = cb2.radius(queries, args.radius, threads)
        del cb2
        rc = rk = rr = True
        for tree in (lbvh, sah):
            for wide in (False, True):
                pb1.closest(tree, wide=wide)
                rc &= bool(np.array_equal(pb1.read(), mv_c))
                pbk.knn(tree, args.k, wide=wide)
                rk &= bool(np.array_equal(pbk.read(), mv_k))
                _, o2, k2 = pb1.radius(tree, args.radius, wide=wide, keep=True)
                rr &= bool(np.array_equal(o2, mv_o) and np.array_equal(k2, mv_keys))
        parity["refit_closest"] = rc
        parity["refit_knn"] = rk
        parity["refit_radius"] = rr
        parity["refit_validate"] = why == "" and why_sah == ""
        refit_roof = refit_roofline(A, lbvh, dirty, n, rt)
        extra["refit"] = {"dirty_tris": int(rt.updated), "upload_ms": round(rt.upload_ms, 4),
                          "mark_ms": round(rt.mark_ms, 4), "merge_ms": round(rt.merge_ms, 4),
                          "wide_ms": round(rt.wide_ms, 4), "total_ms": round(rt.total_ms, 4),
                          "sah_tree_total_ms": round(rt_sah.total_ms, 4), "roofline": refit_roof}
        log(f"refit: {extra['refit']} validate={why!r} {why_sah!r}")
        lbvh.refit(tris)
        sah.refit(tris)
        # full refit from device-resident scene-order items (no upload): every leaf
        # re-gathered and every internal box merged, 2S + 28 per leaf + 84 per node
        full = {}
        for name, tree in (("lbvh", lbvh), ("sah", sah)):
            ms = [tree.refit_device(dtris.ptr.value, n) for _ in range(6)][1:]
            fms = statistics.median(ms)
            alg = n * (2 * TRI_BYTES + 28) + (n - 1) * 84
            full[name] = {"ms": round(fms, 4), "alg_bytes": alg, "achieved": round(alg / (fms * 1e-3) / 1e9, 2),
                          "frac": round(alg / (fms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        extra["refit_full_device"] = full
        parity["refit_full_validate"] = lbvh.validate(tris) == "" and sah.validate(tris) == ""
        log(f"full refit (device-resident items): {full}")
        log("refit back to the original scene done")

        # ---- rays: 1M down rays on the 256 x 256 terrain ----------------------------
        tr1, rays = S.workload_c1(seed=1)
        b1 = A.Bvh(A.PRIM_TRI, tr1)
        b1.build_wide()
        log("rays: tree built")
        rb = A.RayBuffer(rays)
        for _ in range(3):
            rb.cast(b1)
        rms = [rb.cast(b1) for _ in range(10)]
        got, rs = rb.summarize()
        cbr = O.CpuBvh(0, tr1)
        exp_r = cbr.ray(rays, threads)
        parity["rays"] = bool(np.array_equal(got, exp_r))
        ray_ms = statistics.median(rms)
        extra["rays_c1"] = {"mrays_per_s": round(rays.shape[0] / (ray_ms * 1e-3) / 1e6, 2),
                            "ms": round(ray_ms, 4), "hits": int(rs.hits), "sum_t": round(rs.sum_t, 6),
                            "hash": hex(rs.hash),
                            "build_ms": round(b1.timings.total_ms, 4)}
        log(f"rays: {extra['rays_c1']} parity={parity['rays']}")
        del cbr, rb, b1

        # ---- the other workloads: non-uniform scenes, gradients, batches --------------
        if not args.no_extras:
            for name, leg in (("c3", leg_c3), ("c4_grad", leg_c4), ("c5_batched", leg_c5)):
                t0 = time.time()
                try:
                    extra[name], ok = leg(A, S, O, threads)
                    parity[name] = ok
                except Exception as e:  # reported, never hidden: the parity flag fails
                    extra[name] = {"error": f"{type(e).__name__}: {e}"}
                    parity[name] = False
                log(f"{name} ({time.time() - t0:.1f}s): {extra[name]}")

        # ---- CPU baseline on a bounded sample ------------------------------------------
        if not args.no_cpu:
            cpu_baseline = cpu_leg(O, tris, queries, args, threads)
            log(f"cpu baseline: {cpu_baseline}")

    out = {
        "metric": "Mqueries/s (closest + kNN16 + radius r=2, bit-exact), 2M-triangle terrain",
        "value": round(value, 3),
        "unit": "Mqueries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded z-up sinusoid terrain, uniform queries)",
        "config": {"workload": "terrain 1024x1024 cells (2,097,152 tris), 100k queries, closest + "
                               f"kNN{args.k} + radius {args.radius}, Morton order, 4-wide nodes, "
                               f"{args.builder.upper()}-built index",
                   "triangles": n, "queries": nq, "k": args.k, "radius": args.radius,
                   "builder": args.builder, "radius_results": radius_total},
        "other_index": other_index,
        "per_op_ms": {k: round(v, 4) for k, v in per_op.items()},
        "per_op_mqps": {k: round(nq / (v * 1e-3) / 1e6, 2) for k, v in per_op.items()},
        "per_op_kernel_ms": {k: round(v, 4) for k, v in per_kernel.items()},
        "tests_per_query": {k: {"node": round(v[0], 2), "leaf": round(v[1], 2),
                                "opened": round(v[2], 2)} for k, v in work.items()},
        "tests_per_query_sequential_walk": {k: {"node": round(v[0], 2), "leaf": round(v[1], 2),
                                                "opened": round(v[2], 2)} for k, v in seq.items()},
        "binary_kernel_ms": binary,
        "roofline": roofline,
        "build": build,
        "parity": parity,
        "cpu_baseline": cpu_baseline,
        "device": dev,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)
    return 0 if all(parity.values()) else 3


def refit_roofline(A, tree, dirty, n, rt):
    """Bytes of a dirty refit (2S + 28 per dirty leaf + 84 per internal node merged)
    over the refit's device time after the upload.  The internal nodes merged are the
    ancestors of the dirty leaves, counted here from the tree's node array; the path
    is the library's rule: marked ancestors only while at most 1 leaf in 256 is
    dirty, else the whole bottom-up pass (every leaf and internal node)."""
    nodes, ids = tree.read()
    links = nodes.view(np.int32).reshape(-1, 8)[:, 6:8]
    parent = np.full(2 * n - 1, -1, np.int64)
    internal = np.arange(n - 1)
    parent[links[: n - 1, 0]] = internal
    parent[links[: n - 1, 1]] = internal
    slot = np.empty(n, np.int64)
    slot[ids] = np.arange(n)
    seen = np.zeros(n - 1, bool)
    cur = np.unique(parent[n - 1 + slot[np.unique(dirty)]])
    cur = cur[cur >= 0]
    while cur.size:
        cur = cur[~seen[cur]]
        seen[cur] = True
        cur = np.unique(parent[cur])
        cur = cur[cur >= 0]
    marked = int(seen.sum())
    m = int(np.unique(dirty).size)
    dense = m * 256 > n
    alg = m * (2 * TRI_BYTES + 28) + marked * 84
    done = m * (2 * TRI_BYTES + 8) + (n * (TRI_BYTES + 32) + (n - 1) * 84 if dense else marked * 84)
    dev_ms = rt.mark_ms + rt.merge_ms
    return {"path": "dense" if dense else "marked", "dirty_leaves": m, "marked_internal": int(marked),
            "alg_bytes": int(alg), "path_bytes": int(done), "device_ms": round(dev_ms, 4),
            "achieved": round(alg / (dev_ms * 1e-3) / 1e9, 2),
            "frac": round(alg / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "path_achieved": round(done / (dev_ms * 1e-3) / 1e9, 2)}


def _median_ms(fn, reps=5):
    fn()
    return statistics.median(fn() for _ in range(reps))


def leg_c3(A, S, O, threads):
    """Non-uniform scenes: city rays (LBVH and 32-bin SAH), 1M-point kNN8 / radius 1,
    1M-Gaussian closest / kNN8 (each cloud on an LBVH and on a SAH tree); 20,000
    uniform queries per scene."""
    out, ok = {}, True
    city = S.city(131072, 1)
    rays = S.cosine_down_rays(city, 1_000_000, 2)
    want = O.CpuBvh(0, city).ray(rays, threads)
    rb = A.RayBuffer(rays)
    for name, builder in (("lbvh", A.BUILDER_LBVH), ("sah", A.BUILDER_SAH)):
        b = A.Bvh(0, city, builder)
        b.build_wide()
        ms = _median_ms(lambda: rb.cast(b))
        rb.cast(b, stats=True)
        st = A.last_ray_stats()
        got, _ = rb.summarize()
        same = bool(np.array_equal(got, want))
        ok &= same
        out[f"city_rays_{name}"] = {"mrays_per_s": round(rays.shape[0] / ms / 1e3, 1), "ms": round(ms, 4),
                                    "build_ms": round(b.timings.total_ms, 3),
                                    "node_tests_per_ray": round(st.node_tests / rays.shape[0], 2),
                                    "parity": same}
        b.close()
    out["city_triangles"] = int(city.shape[0])
    rng = np.random.default_rng(12)
    clouds = ((1, S.point_cloud(1 << 20, 1), "points"), (2, S.gaussians(1 << 20, 1), "gaussians"))
    q = rng.uniform(0.0, 64.0, (20_000, 3)).astype(np.float32)
    for kind, items, name in clouds:
        ref = O.CpuBvh(kind, items)
        want_c, want_k = ref.knn(q, 0, threads), ref.knn(q, 8, threads)
        want_r = ref.radius(q, 1.0, threads) if kind == 1 else None
        del ref
        for builder, label in ((A.BUILDER_LBVH, name), (A.BUILDER_SAH, name + "_sah")):
            b = A.Bvh(kind, items, builder)
            b.build_wide()
            pb1, pb8 = A.PointBuffer(q, 1), A.PointBuffer(q, 8)
            res = {"build_ms": round(b.timings.total_ms, 3)}
            ms_c = _median_ms(lambda: pb1.closest(b))
            same_c = bool(np.array_equal(pb1.read(), want_c))
            ms_k = _median_ms(lambda: pb8.knn(b, 8))
            same_k = bool(np.array_equal(pb8.read(), want_k))
            res.update({"closest_mq_per_s": round(q.shape[0] / ms_c / 1e3, 2),
                        "knn8_mq_per_s": round(q.shape[0] / ms_k / 1e3, 2), "parity_closest": same_c,
                        "parity_knn8": same_k})
            ok &= same_c and same_k
            if want_r is not None:
                ms_r = _median_ms(lambda: pb1.radius(b, 1.0))
                _, offs, keys = pb1.radius(b, 1.0, keep=True)
                same_r = bool(np.array_equal(offs, want_r[0]) and np.array_equal(keys, want_r[1]))
                res.update({"radius1_mq_per_s": round(q.shape[0] / ms_r / 1e3, 2), "parity_radius": same_r})
                ok &= same_r
            out[label] = res
            b.close()
    return out, bool(ok)


def leg_c4(A, S, O, threads):
    """Differentiable closest-point and ray queries on the 327,680-triangle icosphere:
    forward-with-save + adjoint against the plain forward, through the C ABI."""
    import ctypes

    from oracle import adjoint as J

    L = A.lib()
    V, F, tris = S.icosphere(levels=7, radius=10.0, noise=0.1, seed=1)
    rng = np.random.default_rng(13)
    q = rng.uniform(-15, 15, (262_144, 3)).astype(np.float32)
    rays = S.sphere_rays(1_000_000, seed=14)
    dt, dq, dr = A.DeviceArray(tris), A.DeviceArray(q), A.DeviceArray(rays)
    b = A.Bvh.from_device(0, dt.ptr.value, tris.shape[0])
    b.build_wide()
    nq, nr = q.shape[0], rays.shape[0]
    res = A.DeviceArray.empty((nq, 2), np.uint32)
    aux = A.DeviceArray.empty(nq, np.int32)
    rres = A.DeviceArray.empty(nr, np.uint64)
    hit = A.DeviceArray.empty(nr, np.int32)
    ones_q, ones_r = A.DeviceArray(np.ones(nq, np.float32)), A.DeviceArray(np.ones(nr, np.float32))
    gq = A.DeviceArray.empty((nq, 3), np.float32)
    go, gd = A.DeviceArray.empty((nr, 3), np.float32), A.DeviceArray.empty((nr, 3), np.float32)
    gt = A.DeviceArray(np.zeros(tris.shape, np.float32))
    ms = ctypes.c_double()

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        A.check(L.qr_dev_synchronize())
        return 1000.0 * (time.perf_counter() - t0)

    def plain():
        A.check(L.qr_closest_device(b.h, dq.ptr, nq, res.ptr, A.ORDER_MORTON, 1, ctypes.byref(ms)))
        A.check(L.qr_raycast_device(b.h, dr.ptr, nr, rres.ptr, A.ORDER_MORTON, 1, ctypes.byref(ms)))

    def forward_save():
        A.check(L.qr_closest_device(b.h, dq.ptr, nq, res.ptr, A.ORDER_MORTON, 1, ctypes.byref(ms)))
        A.check(L.qr_closest_save_aux(b.h, dq.ptr, nq, res.ptr, aux.ptr))
        A.check(L.qr_raycast_save_device(b.h, dr.ptr, nr, rres.ptr, hit.ptr, A.ORDER_MORTON, 1, ctypes.byref(ms)))

    def backward():
        A.check(L.qr_adjoint_closest(b.h, dq.ptr, nq, res.ptr, aux.ptr, ones_q.ptr, gq.ptr, gt.ptr))
        A.check(L.qr_adjoint_ray(b.h, dr.ptr, nr, rres.ptr, hit.ptr, ones_r.ptr, go.ptr, gd.ptr, gt.ptr))

    t_plain = _median_ms(lambda: timed(plain))
    t_fwd = _median_ms(lambda: timed(forward_save))
    t_bwd = _median_ms(lambda: timed(backward))
    plain()
    plain_res = res.download()
    forward_save()
    same_fwd = bool(np.array_equal(res.download(), plain_res))
    want = O.CpuBvh(0, tris).knn(q[:20000], 0, threads)
    same_fwd &= bool(np.array_equal(plain_res[:20000], want[:, 0, :]))
    # gradient spot check against the FP64 CPU adjoint on the first 4,096 queries
    A.check(L.qr_adjoint_closest(b.h, dq.ptr, nq, res.ptr, aux.ptr, ones_q.ptr, gq.ptr, None))
    prim = plain_res[:, 1].view(np.int32)
    cq, _ = J.tri_adjoint(q[:4096].astype(np.float64), tris.astype(np.float64), prim[:4096], np.ones(4096))
    g = gq.download()[:4096].astype(np.float64)
    floor = 0.05 * float(np.sqrt(np.mean(cq ** 2)))
    err = float(np.max(np.abs(g - cq) / np.maximum(np.abs(cq), floor)))
    out = {"triangles": int(tris.shape[0]), "queries": nq, "rays": nr,
           "plain_forward_ms": round(t_plain, 3), "forward_with_save_ms": round(t_fwd, 3),
           "adjoint_ms": round(t_bwd, 3),
           "grad_step_vs_two_plain_passes": round((t_fwd + t_bwd) / (2.0 * t_plain), 3),
           "forward_bit_exact": same_fwd, "adjoint_max_rel_err_vs_fp64": err}
    del b, dt, dq, dr, res, aux, rres, hit, gq, go, gd, gt
    gauss, ok_g = leg_c4_gaussians(A, S, O, J, threads, timed)
    out["gaussians"] = gauss
    return out, bool(same_fwd and err < 1e-4 and ok_g)


def leg_c4_gaussians(A, S, O, J, threads, timed):
    """The differentiable closest-point pass on the 1,048,576-Gaussian scene (the
    non-uniform scenes' Gaussians) with 262,144 uniform queries: plain forward,
    forward with the saved Eberly roots, and the adjoint into query and Gaussian
    (mean, quaternion, scale) gradients."""
    import ctypes

    L = A.lib()
    g = S.gaussians(1 << 20, 1)
    q = np.random.default_rng(21).uniform(0.0, 64.0, (262_144, 3)).astype(np.float32)
    nq, ng = q.shape[0], g.shape[0]
    dg, dq = A.DeviceArray(g), A.DeviceArray(q)
    b = A.Bvh.from_device(2, dg.ptr.value, ng)
    b.build_wide()
    res = A.DeviceArray.empty((nq, 2), np.uint32)
    aux = A.DeviceArray.empty(nq, np.int32)
    ones = A.DeviceArray(np.ones(nq, np.float32))
    gq = A.DeviceArray.empty((nq, 3), np.float32)
    gg = A.DeviceArray(np.zeros((ng, 10), np.float32))
    ms = ctypes.c_double()

    def plain():
        A.check(L.qr_closest_device(b.h, dq.ptr, nq, res.ptr, A.ORDER_MORTON, 1, ctypes.byref(ms)))

    def forward_save():
        plain()
        A.check(L.qr_closest_save_aux(b.h, dq.ptr, nq, res.ptr, aux.ptr))

    def backward():
        A.check(L.qr_adjoint_closest(b.h, dq.ptr, nq, res.ptr, aux.ptr, ones.ptr, gq.ptr, gg.ptr))

    t_plain = _median_ms(lambda: timed(plain))
    t_fwd = _median_ms(lambda: timed(forward_save))
    t_bwd = _median_ms(lambda: timed(backward))
    plain()
    plain_res = res.download()
    forward_save()
    same = bool(np.array_equal(res.download(), plain_res))
    want = O.CpuBvh(2, g).knn(q[:20000], 0, threads)
    same &= bool(np.array_equal(plain_res[:20000], want[:, 0, :]))
    # gradients of the first 4,096 queries' loss against the FP64 CPU adjoint
    m = 4096
    gg0 = A.DeviceArray(np.zeros((ng, 10), np.float32))
    A.check(L.qr_adjoint_closest(b.h, dq.ptr, m, res.ptr, aux.ptr, ones.ptr, gq.ptr, gg0.ptr))
    prim = plain_res[:m, 1].view(np.int32)
    cq, cg = J.gauss_adjoint(q[:m].astype(np.float64), g.astype(np.float64), prim, np.ones(m))
    got_q, got_g = gq.download()[:m].astype(np.float64), gg0.download().astype(np.float64)
    err = 0.0
    for got, ref in ((got_q, cq), (got_g[:, 0:3], cg[:, 0:3]), (got_g[:, 3:7], cg[:, 3:7]),
                     (got_g[:, 7:10], cg[:, 7:10])):
        floor = 0.05 * float(np.sqrt(np.mean(ref[ref != 0] ** 2)))
        err = max(err, float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), floor))))
    out = {"gaussians": ng, "queries": nq, "plain_forward_ms": round(t_plain, 3),
           "forward_with_save_ms": round(t_fwd, 3), "adjoint_ms": round(t_bwd, 3),
           "grad_step_vs_two_plain_passes": round((t_fwd + t_bwd) / (2.0 * t_plain), 3),
           "forward_bit_exact": same, "adjoint_max_rel_err_vs_fp64": err}
    b.close()
    return out, bool(same and err < 1e-4)


def leg_c5(A, S, O, threads):
    """4,096 uneven scenes: one segmented build + persistent-kernel queries (hipGraph,
    autotuned) against the looped plain per-scene path; results compared whole."""
    import ctypes

    L = A.lib()
    items, poff, cp, coff, kp, koff = S.workload_c5(seed=1)
    S_ = poff.shape[0] - 1
    di = A.DeviceArray(items)
    builds = []
    A.check(L.qr_set_build_phase_timing(0))  # (as for the plain builds: no phase events)
    try:
        for _ in range(3):
            bt = A.Batch.from_device(0, di.ptr.value, poff)
            builds.append(bt.timings.total_ms)
            if len(builds) < 3:
                bt.close()
    finally:
        A.check(L.qr_set_build_phase_timing(1))
    plan_c = A.BatchPlan(bt, cp, coff, 0, tune=True)
    plan_k = A.BatchPlan(bt, kp, koff, 8, tune=True)
    ms_c = _median_ms(lambda: plan_c.run())
    ms_k = _median_ms(lambda: plan_k.run())
    got_c, got_k = plan_c.read(), plan_k.read()
    nq = int(coff[-1] + koff[-1])
    # looped plain path: one build and two query calls per scene, device pointers
    dc, dk = plan_c.points, plan_k.points
    oc = A.DeviceArray.empty((int(coff[-1]), 2), np.uint32)
    ok_ = A.DeviceArray.empty((int(koff[-1]), 8, 2), np.uint32)
    ms = ctypes.c_double()
    trees = []
    t0 = time.perf_counter()
    for s in range(S_):
        trees.append(A.Bvh.from_device(0, di.ptr.value + int(poff[s]) * 36, int(poff[s + 1] - poff[s])))
    A.check(L.qr_dev_synchronize())
    t_build = 1000.0 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    for s in range(S_):
        h = trees[s].h
        n1, n2 = int(coff[s + 1] - coff[s]), int(koff[s + 1] - koff[s])
        if n1:
            A.check(L.qr_closest_device(h, dc.at(int(coff[s]) * 12), n1, oc.at(int(coff[s]) * 8),
                                        A.ORDER_MORTON, 0, ctypes.byref(ms)))
        if n2:
            A.check(L.qr_knn_device(h, dk.at(int(koff[s]) * 12), n2, 8, ok_.at(int(koff[s]) * 64),
                                    A.ORDER_MORTON, 0, ctypes.byref(ms)))
    A.check(L.qr_dev_synchronize())
    t_query = 1000.0 * (time.perf_counter() - t0)
    loop_c, loop_k = oc.download(), ok_.download()
    same_c = bool(np.array_equal(got_c.reshape(-1, 2), loop_c))
    same_k = bool(np.array_equal(got_k, loop_k))
    hashes_equal = sum(int(O.knn_hash(got_k[koff[s]:koff[s + 1]], 8)[0] ==
                           O.knn_hash(loop_k[koff[s]:koff[s + 1]], 8)[0]) for s in range(S_))
    trees_equal = sum(int(bt.tree_hash(s) == trees[s].tree_hash()) for s in range(0, S_, 64))
    for tr in trees:
        tr.close()
    bb_ms = statistics.median(builds)
    bb_alg = int(poff[-1]) * (5 * TRI_BYTES + 232)
    out = {"scenes": S_, "triangles": int(poff[-1]), "queries": nq,
           "batched_build_ms": round(bb_ms, 3),
           "batched_build_roofline": {"bound": "hbm", "alg_bytes": bb_alg,
                                      "achieved": round(bb_alg / (bb_ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                                      "unit": "GB/s", "frac": round(bb_alg / (bb_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
           "batched_query_ms": round(ms_c + ms_k, 3),
           "batched_mq_per_s": round(nq / (ms_c + ms_k) / 1e3, 2),
           "plain_loop_build_ms": round(t_build, 1), "plain_loop_query_ms": round(t_query, 1),
           "plain_loop_mq_per_s": round(nq / t_query / 1e3, 2),
           "speedup_queries": round(t_query / (ms_c + ms_k), 1),
           "speedup_build_plus_queries": round((t_build + t_query) / (statistics.median(builds) + ms_c + ms_k), 1),
           "tuned": {"closest": plan_c.config(), "knn8": plan_k.config()},
           "results_equal": {"closest": same_c, "knn8": same_k},
           "knn8_fingerprints_equal": hashes_equal, "tree_hashes_equal_sampled": trees_equal}
    plan_c.close()
    plan_k.close()
    bt.close()
    return out, bool(same_c and same_k)


def calibration_copy(A, mib=256):
    """One streaming device copy of known size (mib MiB read, mib MiB written), so
    that a counter run can scale FETCH_SIZE / WRITE_SIZE to bytes."""
    import ctypes

    nbytes = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    A.check(A.lib().qr_dev_alloc(nbytes, ctypes.byref(src)))
    A.check(A.lib().qr_dev_alloc(nbytes, ctypes.byref(dst)))
    ms = ctypes.c_double()
    for _ in range(2):
        A.check(A.lib().qr_dev_copy(dst, src, nbytes, ctypes.byref(ms)))
    log(f"calibration copy: {mib} MiB in {ms.value:.3f} ms ({2 * nbytes / ms.value / 1e6:.0f} GB/s)")
    A.lib().qr_dev_free(src)
    A.lib().qr_dev_free(dst)


def traffic_for(kernel, args):
    """HBM bytes per launch of the dominant traversal kernel, measured offline with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this exact workload and scaled by the
    same counters' reading for a copy of known size; the figures are kept in
    profiles/hbm_traffic.txt.  None when no measurement for this configuration exists."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    key = (f"c2 cells={args.cells} queries={args.queries} k={args.k} radius={args.radius} "
           f"builder={args.builder} kernel={kernel}")
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith(key + " "):
                    m = re.search(r"\bbytes=([0-9.eE+]+)", ln[len(key):])
                    return int(float(m.group(1))) if m else None
    except OSError:
        pass
    return None


def cpu_leg(O, tris, queries, args, threads):
    """Times the reference CPU path (oracle/_ref) -- or the oracle restatement when
    the reference build is absent -- on a bounded sample of the same queries."""
    ref = O.ref() if O.ref_available() else None
    kind = "reference" if ref is not None else "port"
    p = np.ascontiguousarray(queries, dtype=np.float32)
    if ref is not None:
        f32p, u32p, i32p = O._f32p, O._u32p, O._i32p
        t0 = time.time()
        h = ref.ref_bvh_new(np.ascontiguousarray(tris).ctypes.data_as(f32p), tris.shape[0])
        build_s = time.time() - t0

        def run(m, thr):
            q = p[:m]
            out1 = np.zeros((m, 1, 2), np.uint32)
            outk = np.zeros((m, args.k, 2), np.uint32)
            cnt = np.zeros(m, np.int32)
            ref.ref_bvh_knn(h, q.ctypes.data_as(f32p), m, 0, out1.ctypes.data_as(u32p), thr)
            ref.ref_bvh_knn(h, q.ctypes.data_as(f32p), m, args.k, outk.ctypes.data_as(u32p), thr)
            ref.ref_bvh_radius(h, q.ctypes.data_as(f32p), m, args.radius, cnt.ctypes.data_as(i32p), None, None, thr)
    else:
        t0 = time.time()
        cb = O.CpuBvh(0, tris)
        build_s = time.time() - t0

        def run(m, thr):
            q = p[:m]
            cb.knn(q, 0, thr)
            cb.knn(q, args.k, thr)
            cb.radius(q, args.radius, thr)

    # calibrate the single-thread sample to the time budget
    m = 500
    t0 = time.time()
    run(m, 1)
    dt = max(time.time() - t0, 1e-3)
    m1 = int(min(p.shape[0], max(500, m * (0.5 * args.cpu_seconds) / dt)))
    t0 = time.time()
    run(m1, 1)
    single = 3 * m1 / (time.time() - t0) / 1e6
    m2 = int(min(p.shape[0], m1 * threads))
    t0 = time.time()
    run(m2, threads)
    multi = 3 * m2 / (time.time() - t0) / 1e6
    if ref is not None:
        ref.ref_bvh_free(h)
    try:
        import subprocess
        model = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        model = next((ln.split(":", 1)[1].strip() for ln in model.splitlines()
                      if ln.startswith("Model name")), "unknown")
    except Exception:
        model = "unknown"
    return {"value": round(multi, 4), "unit": "Mqueries/s", "cores": threads, "kind": kind,
            "single_thread_value": round(single, 4),
            "sample": f"closest + kNN{args.k} + radius {args.radius} over the first {m1} queries "
                      f"(1 thread) and {m2} queries ({threads} threads) of the same workload; "
                      f"CPU BVH build {build_s:.2f}s not counted",
            "cpu_model": model, "nproc": os.cpu_count()}


if __name__ == "__main__":
    sys.exit(main())
