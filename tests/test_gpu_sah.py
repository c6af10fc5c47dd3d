"""Binned-SAH builder (32 bins): structure, determinism and bit-exact queries.

Query results do not depend on the tree (conservative pruning, (dist2, index)
ordering), so every SAH result set must equal the brute-force oracle exactly, as
the reference requires of its own SAH path (python/test_quarry.py:241-257,
docs/baseline.md:843-844).  The node bytes of a 32-bin SAH have no reference
counterpart (the reference bins 16 ways); the tree is pinned by validate(), by
determinism across builds and by the query results.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def eq(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


@pytest.fixture(scope="module")
def A(qabi):
    if qabi.describe() is None:
        pytest.fail("no HIP device visible to the GPU tests")
    return qabi


def _items(kind, n, seed, scenes):
    rng = np.random.default_rng(seed)
    if kind == 0:
        a = rng.uniform(-30, 30, (n, 3))
        return np.concatenate([a, a + rng.normal(0, 1.5, (n, 3)), a + rng.normal(0, 1.5, (n, 3))], 1).astype(np.float32)
    if kind == 1:
        return scenes.point_cloud(n, seed)
    return scenes.gaussians(n, seed)


@pytest.mark.parametrize("kind,n", [(0, 2), (0, 3), (0, 64), (0, 65), (0, 4097), (0, 30000), (1, 2),
                                    (1, 20000), (2, 9000)])
def test_sah_structure(A, scenes, kind, n):
    print(f"sah structure kind={kind} n={n}", flush=True)
    items = _items(kind, n, n + 17 * kind, scenes)
    b = A.Bvh(kind, items, A.BUILDER_SAH)
    inf = b.info()
    assert inf.builder == A.BUILDER_SAH
    assert inf.node_count == 2 * n - 1
    assert b.validate(items) == ""
    assert 1 <= inf.depth < 64 if n > 1 else True
    nodes, ids = b.read()
    assert eq(np.sort(ids), np.arange(n))
    b2 = A.Bvh(kind, items, A.BUILDER_SAH)
    assert b2.tree_hash() == b.tree_hash(), "SAH build is not deterministic"


def test_sah_duplicates_and_degenerate(A):
    # all centroids equal (every axis degenerate -> count/2 splits) and a
    # two-valued cloud (planes exist only on one axis)
    pts = np.zeros((5000, 3), np.float32)
    b = A.Bvh(1, pts, A.BUILDER_SAH)
    assert b.validate(pts) == ""
    assert b.info().depth <= 13
    pts2 = np.zeros((3000, 3), np.float32)
    pts2[1500:, 0] = 1.0
    b = A.Bvh(1, pts2, A.BUILDER_SAH)
    assert b.validate(pts2) == ""
    q = np.random.default_rng(2).uniform(-1, 2, (400, 3)).astype(np.float32)
    import oracle as O

    assert eq(b.knn(q, 8)[0], O.knn_brute(1, pts2, q, 8, threads=THREADS))


def test_sah_queries_vs_brute(A, orc, scenes):
    rng = np.random.default_rng(3)
    tris = np.concatenate([scenes.terrain(24, 5), _items(0, 900, 4, scenes) * np.float32(0.4)])
    q = rng.uniform(-16, 30, (1500, 3)).astype(np.float32)
    b = A.Bvh(0, tris, A.BUILDER_SAH)
    b.build_wide()
    want_c = orc.knn_brute(0, tris, q, 0, threads=THREADS)
    want_k = orc.knn_brute(0, tris, q, 16, threads=THREADS)
    want_r = orc.radius_brute(0, tris, q, 2.0, threads=THREADS)
    rays = scenes.cosine_down_rays(tris, 3000, 5)
    want_ray = orc.ray_brute(tris, rays, threads=THREADS)
    for wide in (False, True):
        for order in (A.ORDER_INPUT, A.ORDER_MORTON):
            assert eq(b.closest(q, order=order, wide=wide)[0], want_c)
            assert eq(b.knn(q, 16, order=order, wide=wide)[0], want_k)
            o, k, _, _ = b.radius(q, 2.0, order=order, wide=wide)
            assert eq(o, want_r[0]) and eq(k, want_r[1])
            assert eq(b.raycast(rays, order=order, wide=wide)[0], want_ray)


@pytest.mark.parametrize("kind", [1, 2])
def test_sah_points_gaussians_vs_brute(A, orc, scenes, kind):
    items = _items(kind, 12000, 21 + kind, scenes)
    q = np.random.default_rng(6).uniform(-4, 68, (900, 3)).astype(np.float32)
    b = A.Bvh(kind, items, A.BUILDER_SAH)
    b.build_wide()
    for wide in (False, True):
        assert eq(b.knn(q, 8, order=A.ORDER_MORTON, wide=wide)[0], orc.knn_brute(kind, items, q, 8, threads=THREADS))
        o, k, _, _ = b.radius(q, 1.0, wide=wide)
        wo, wk = orc.radius_brute(kind, items, q, 1.0, threads=THREADS)
        assert eq(o, wo) and eq(k, wk)


def test_sah_refit(A, orc, scenes):
    tris = scenes.terrain(48, 2)
    b = A.Bvh(0, tris, A.BUILDER_SAH)
    b.build_wide()
    moved = tris.copy()
    dirty = np.arange(0, tris.shape[0], 37, dtype=np.int32)
    moved[dirty] += np.float32(0.3)
    b.refit(moved, dirty)
    assert b.validate(moved) == ""
    q = np.random.default_rng(8).uniform(-2, 50, (1000, 3)).astype(np.float32)
    assert eq(b.closest(q, wide=True)[0], orc.knn_brute(0, moved, q, 0, threads=THREADS))


def test_sah_city_fewer_node_tests(A, orc, scenes):
    """On the non-uniform city the SAH tree needs fewer node tests per ray than the
    LBVH (the reference reports 28 vs 173, docs/baseline.md:1073-1077)."""
    tris = scenes.city(4096, 3)
    rays = scenes.cosine_down_rays(tris, 20000, 4)
    want = orc.CpuBvh(0, tris).ray(rays, threads=THREADS)
    tests = {}
    for builder in (A.BUILDER_LBVH, A.BUILDER_SAH):
        b = A.Bvh(0, tris, builder)
        got, _, st = b.raycast(rays, stats=True)
        assert eq(got, want), builder
        tests[builder] = st.node_tests / rays.shape[0]
    print("node tests per ray:", tests, flush=True)
    assert tests[A.BUILDER_SAH] < tests[A.BUILDER_LBVH]


def test_build_phase_timing_levels(A, scenes):
    """qr_set_build_phase_timing: 0 fills total_ms only, 1 (the default) the phase
    fields too, 2 also the SAH per-level split; the trees are the same at every level."""
    L = A.lib()
    items = _items(0, 300000, 11, scenes)
    hashes = {A.BUILDER_LBVH: set(), A.BUILDER_SAH: set()}
    try:
        for level in (0, 1, 2):
            A.check(L.qr_set_build_phase_timing(level))
            for builder in (A.BUILDER_LBVH, A.BUILDER_SAH):
                print(f"phase timing level={level} builder={builder}", flush=True)
                b = A.Bvh(0, items, builder)
                t = b.timings
                assert t.total_ms > 0.0
                if level == 0:
                    assert t.morton_ms == 0.0 and t.refit_ms == 0.0
                else:
                    assert t.morton_ms > 0.0 and t.refit_ms > 0.0
                if builder == A.BUILDER_SAH:
                    assert (t.sah_bin_ms > 0.0) == (level == 2)
                hashes[builder].add(b.tree_hash())
                b.close()
        with pytest.raises(Exception):
            A.check(L.qr_set_build_phase_timing(3))
    finally:
        A.check(L.qr_set_build_phase_timing(1))
    assert all(len(h) == 1 for h in hashes.values())
