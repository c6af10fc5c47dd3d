"""Batched mode: many uneven scenes built and queried together.

Every scene's tree must equal the plain per-scene build byte for byte (tree hash
and node array), and every query result must equal the plain per-scene query --
closest and kNN k=8, on every kernel variant the autotuner can pick, through the
captured hipGraph over several replays (with eager runs and host work between
them) and directly.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def A(qabi):
    if qabi.describe() is None:
        pytest.fail("no HIP device visible to the GPU tests")
    return qabi


@pytest.fixture(scope="module")
def work(scenes):
    return scenes.workload_c5(seed=3, scenes=40, nmin=256, nmax=6000, qmin=64, qmax=900)


def test_batch_trees_equal_plain(A, work):
    items, poff, *_ = work
    b = A.Batch(A.PRIM_TRI, items, poff)
    S = poff.shape[0] - 1
    assert b.info()[0] == S
    for s in range(S):
        part = items[poff[s]:poff[s + 1]]
        plain = A.Bvh(A.PRIM_TRI, part)
        nodes, ids = b.read_scene(s)
        pn, pid = plain.read()
        assert np.array_equal(nodes, pn), s
        assert np.array_equal(ids, pid), s
        assert b.tree_hash(s) == plain.tree_hash()


@pytest.mark.parametrize("cap", [32, 5, 0])
def test_batch_trees_equal_plain_large_scenes(A, scenes, cap):
    """Scenes from 3 to 70,001 primitives in one batch (one to three hand-up levels of
    the per-scene merge, and windows of one leaf), at the default frontier capacity and
    with units sent through the device-wide arrival path (cap 5 and 0): every tree
    equals the plain per-scene build."""
    rng = np.random.default_rng(23 + cap)
    sizes = [70001, 3, 20000, 257, 9000, 513]
    parts = [rng.uniform(-9, 9, (n, 3)).astype(np.float32) for n in sizes]
    poff = np.zeros(len(parts) + 1, np.int32)
    np.cumsum(sizes, out=poff[1:])
    try:
        A.check(A.lib().qr_set_frontier_cap(cap))
        b = A.Batch(1, np.concatenate(parts), poff)
        for s, part in enumerate(parts):
            plain = A.Bvh(1, part)
            nodes, ids = b.read_scene(s)
            pn, pid = plain.read()
            assert np.array_equal(nodes, pn), (cap, s)
            assert np.array_equal(ids, pid), (cap, s)
    finally:
        A.check(A.lib().qr_set_frontier_cap(32))


def _plain_results(A, items, poff, pts, qoff, k):
    outs = []
    for s in range(poff.shape[0] - 1):
        bvh = A.Bvh(A.PRIM_TRI, items[poff[s]:poff[s + 1]])
        q = pts[qoff[s]:qoff[s + 1]]
        if q.shape[0] == 0:
            continue
        outs.append(bvh.closest(q)[0] if k == 0 else bvh.knn(q, k)[0])
    return np.concatenate(outs)


@pytest.mark.parametrize("k", [0, 8])
def test_batch_queries_equal_plain(A, orc, work, k):
    items, poff, cp, coff, kp, koff = work
    pts, qoff = (cp, coff) if k == 0 else (kp, koff)
    b = A.Batch(A.PRIM_TRI, items, poff)
    plan = A.BatchPlan(b, pts, qoff, k, tune=True)
    print("tuned:", plan.config(), flush=True)
    want = _plain_results(A, items, poff, pts, qoff, k)
    for rep in range(3):
        plan.run(graph=True)
        assert np.array_equal(plan.read(), want), ("graph replay", rep)
        plan.run(graph=False)
        assert np.array_equal(plan.read(), want), ("eager", rep)
    # brute force on a few scenes (scene-local indices)
    got = plan.read()
    for s in (0, 1, 7):
        part = items[poff[s]:poff[s + 1]]
        q = pts[qoff[s]:qoff[s + 1]]
        ref = orc.knn_brute(0, part, q, k, threads=THREADS)
        assert np.array_equal(got[qoff[s]:qoff[s + 1]], ref), s


@pytest.mark.parametrize("k", [0, 8, 32])
def test_batch_every_variant(A, work, k):
    """Every configuration the autotuner can pick -- binary or 4-wide nodes, the stack
    in LDS, scratch or registers, with and without Morton order -- gives the plain
    results."""
    items, poff, cp, coff, kp, koff = work
    pts, qoff = (cp, coff) if k == 0 else (kp, koff)
    b = A.Batch(A.PRIM_TRI, items, poff)
    want = _plain_results(A, items, poff, pts, qoff, k)
    for v in range(len(A.BATCH_VARIANTS)):
        for m in (False, True):
            print(f"batch variant {A.BATCH_VARIANTS[v]} morton={m} k={k}", flush=True)
            plan = A.BatchPlan(b, pts, qoff, k, tune=(v, m))
            assert plan.config()["variant"] == A.BATCH_VARIANTS[v]
            plan.run()
            assert np.array_equal(plan.read(), want), (v, m)


def test_batch_wide_deep_trees(A, orc, scenes):
    """4-wide batched walks on scenes whose trees are deep (duplicate items: every
    scene's tree holds long index-ordered chains) and on two-primitive scenes."""
    rng = np.random.default_rng(17)
    parts = [np.repeat(rng.uniform(-5, 5, (40, 3)), 60, axis=0).astype(np.float32), rng.uniform(-5, 5, (2, 3)),
             scenes.point_cloud(5000, 3), rng.uniform(-5, 5, (3, 3))]
    parts = [p.astype(np.float32) for p in parts]
    poff = np.zeros(len(parts) + 1, np.int32)
    np.cumsum([p.shape[0] for p in parts], out=poff[1:])
    qs = [rng.uniform(-6, 6, (200, 3)).astype(np.float32) for _ in parts]
    qoff = np.zeros(len(qs) + 1, np.int32)
    np.cumsum([q.shape[0] for q in qs], out=qoff[1:])
    b = A.Batch(1, np.concatenate(parts), poff)
    for v in (4, 7, 3):
        plan = A.BatchPlan(b, np.concatenate(qs), qoff, 16, tune=(v, True))
        plan.run()
        got = plan.read()
        for s in range(len(parts)):
            ref = orc.knn_brute(1, parts[s], qs[s], 16, threads=THREADS)
            assert np.array_equal(got[qoff[s]:qoff[s + 1]], ref), (v, s)


@pytest.mark.parametrize("kind", [1, 2])
def test_batch_points_and_gaussians(A, orc, scenes, kind):
    rng = np.random.default_rng(4 + kind)
    sizes = rng.integers(300, 3000, 12)
    parts = [scenes.point_cloud(int(n), 10 + i) if kind == 1 else scenes.gaussians(int(n), 10 + i, side=20.0)
             for i, n in enumerate(sizes)]
    poff = np.zeros(len(parts) + 1, np.int32)
    np.cumsum([p.shape[0] for p in parts], out=poff[1:])
    items = np.concatenate(parts)
    qs = [rng.uniform(-2, 66 if kind == 1 else 22, (int(m), 3)).astype(np.float32)
          for m in rng.integers(10, 300, len(parts))]
    qoff = np.zeros(len(qs) + 1, np.int32)
    np.cumsum([q.shape[0] for q in qs], out=qoff[1:])
    pts = np.concatenate(qs)
    b = A.Batch(kind, items, poff)
    for k in (0, 4):
        plan = A.BatchPlan(b, pts, qoff, k, tune=False)
        plan.run()
        got = plan.read()
        for s in range(len(parts)):
            ref = orc.knn_brute(kind, parts[s], qs[s], k, threads=THREADS)
            assert np.array_equal(got[qoff[s]:qoff[s + 1]], ref), (kind, k, s)


def test_batch_contracts(A):
    tris = np.random.default_rng(1).uniform(-1, 1, (10, 9)).astype(np.float32)
    with pytest.raises(A.QuarryError):
        A.Batch(A.PRIM_TRI, tris, np.array([0, 1, 10], np.int32))  # a scene with one primitive
    with pytest.raises(A.QuarryError):
        A.Batch(A.PRIM_TRI, tris, np.array([1, 10], np.int32))
