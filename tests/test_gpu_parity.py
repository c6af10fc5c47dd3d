"""Bit-exact parity of the HIP path with the CPU oracle (through the C ABI).

Every result set is compared whole: ray results as (hit << 32 | t bits), closest /
kNN records as (dist2 bits, index), radius CSR offsets and keys.  The oracle is
brute force for small inputs and the median-split CPU BVH (itself equal to brute
force, tests/test_oracle_*.py) at benchmark sizes.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
THREADS = min(16, os.cpu_count() or 1)


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def A(qabi):
    if qabi.describe() is None:
        pytest.fail("no HIP device visible to the GPU tests")
    return qabi


def eq(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


def random_tris(rng, n, spread=30.0, size=1.5):
    a = rng.uniform(-spread, spread, (n, 3))
    return np.concatenate([a, a + rng.normal(0, size, (n, 3)), a + rng.normal(0, size, (n, 3))], 1).astype(np.float32)


def random_gaussians(rng, n, spread=30.0):
    return np.concatenate([rng.uniform(-spread, spread, (n, 3)), rng.normal(size=(n, 4)),
                           np.exp(rng.uniform(np.log(0.05), np.log(1.5), (n, 3)))], 1).astype(np.float32)


# ---- build --------------------------------------------------------------------------

# (1, 2**21 + 12345): the sort's 8,192-key tiles (from 2M keys on), with a partial
# last tile and many equal Morton codes ordered by index
@pytest.mark.parametrize("kind,n", [(0, 2), (0, 3), (0, 5000), (1, 7000), (2, 3000), (1, (1 << 21) + 12345)])
def test_node_bytes_match_cpu_lbvh(A, orc, scenes, kind, n):
    log(f"lbvh kind={kind} n={n}")
    rng = np.random.default_rng(n + kind)
    items = {0: lambda: random_tris(rng, n), 1: lambda: rng.uniform(-9, 9, (n, 3)).astype(np.float32),
             2: lambda: random_gaussians(rng, n)}[kind]()
    b = A.Bvh(kind, items)
    nodes, ids = b.read()
    cn, cids, depth = orc.lbvh_nodes(kind, items)
    assert eq(ids, cids)
    assert eq(nodes, cn)
    assert b.tree_hash() == orc.fnv_bytes(cn)
    assert b.info().depth == depth
    assert b.validate(items) == ""


def test_terrain_build_deterministic(A, orc, scenes):
    tris = scenes.terrain(256, 1)
    h = {A.Bvh(0, tris).tree_hash() for _ in range(3)}
    assert len(h) == 1
    cn, _, _ = orc.lbvh_nodes(0, tris)
    assert h.pop() == orc.fnv_bytes(cn)


@pytest.mark.parametrize("cap", [0, 1, 5, 32])
def test_lbvh_frontier_cap(A, orc, scenes, cap):
    """The build hands each workgroup's unfinished nodes up through groups of
    workgroups; units with more than `cap` of them take the device-wide arrival
    protocol instead (cap 0: every one).  The tree must not depend on the mix."""
    log(f"frontier cap {cap}")
    rng = np.random.default_rng(40 + cap)
    cases = [(0, scenes.terrain(256, 2)),                                 # 131,072 triangles
             (1, rng.uniform(-9, 9, (70001, 3)).astype(np.float32)),
             (1, np.repeat(rng.uniform(-1, 1, (300, 3)), 37, axis=0).astype(np.float32))]  # duplicates
    try:
        A.check(A.lib().qr_set_frontier_cap(cap))
        for kind, items in cases:
            b = A.Bvh(kind, items)
            cn, cids, depth = orc.lbvh_nodes(kind, items)
            nodes, ids = b.read()
            assert eq(ids, cids) and eq(nodes, cn), (kind, items.shape)
            assert b.info().depth == depth
    finally:
        A.check(A.lib().qr_set_frontier_cap(32))
    with pytest.raises(A.QuarryError):
        A.check(A.lib().qr_set_frontier_cap(33))


def test_duplicate_centroids(A, orc):
    pts = np.zeros((4096, 3), np.float32)  # one Morton cell: order by index only
    pts[::3] = 1.0
    b = A.Bvh(1, pts)
    cn, _, _ = orc.lbvh_nodes(1, pts)
    assert eq(b.read()[0], cn)
    q = np.array([[0.5, 0.5, 0.5], [1, 1, 1], [0, 0, 0]], np.float32)
    assert eq(b.knn(q, 32)[0], orc.knn_brute(1, pts, q, 32))


# ---- rays ---------------------------------------------------------------------------

def test_rays_small_vs_brute(A, orc, scenes):
    rng = np.random.default_rng(7)
    tris = np.concatenate([scenes.terrain(32, 2), random_tris(rng, 500, 16, 2.0) + np.float32(16)])
    b = A.Bvh(0, tris)
    b.build_wide()
    down = scenes.down_rays(tris, 4000, 3)
    o = rng.uniform(-40, 80, (4000, 3))
    d = rng.normal(size=(4000, 3))
    d[:50, :2] = 0.0  # exercise the |d| < 1e-20 axes
    rand = np.concatenate([o, d, np.full((4000, 1), 1e-6), np.full((4000, 1), np.inf)], 1).astype(np.float32)
    rand[100:200, 7] = 5.0  # finite tMax
    for rays in (down, rand):
        want = orc.ray_brute(tris, rays, threads=THREADS)
        for wide in (False, True):
            for order in (A.ORDER_INPUT, A.ORDER_MORTON, A.ORDER_MORTON_DIRECTION):
                got, _, st = b.raycast(rays, order=order, wide=wide, stats=True)
                assert eq(got, want), (wide, order)
                assert st.stack_overflows == 0


def test_rays_c1_full_size(A, orc, scenes):
    tris, rays = scenes.workload_c1(seed=1)
    b = A.Bvh(0, tris)
    b.build_wide()
    want = orc.CpuBvh(0, tris).ray(rays, threads=THREADS)
    for wide in (False, True):
        got, ms, _ = b.raycast(rays, order=A.ORDER_MORTON, wide=wide)
        assert eq(got, want), wide
    log(f"C1 rays: hash {hex(orc.ray_hash(want)[0])}")


# ---- point queries --------------------------------------------------------------------

def check_queries(A, orc, kind, items, q, radius, ks=(1, 4, 8, 16, 32), brute=True):
    b = A.Bvh(kind, items)
    b.build_wide()
    ref = orc.CpuBvh(kind, items)
    want_c = orc.knn_brute(kind, items, q, 0, threads=THREADS) if brute else ref.knn(q, 0, THREADS)
    want_r = orc.radius_brute(kind, items, q, radius, threads=THREADS) if brute else ref.radius(q, radius, THREADS)
    for wide in (False, True):
        for order in (A.ORDER_INPUT, A.ORDER_MORTON):
            assert eq(b.closest(q, order=order, wide=wide)[0], want_c), ("closest", wide, order)
            o, k, _, _ = b.radius(q, radius, order=order, wide=wide)
            assert eq(o, want_r[0]) and eq(k, want_r[1]), ("radius", wide, order)
    for k in ks:
        want_k = orc.knn_brute(kind, items, q, k, threads=THREADS) if brute else ref.knn(q, k, THREADS)
        for wide in (False, True):
            assert eq(b.knn(q, k, order=A.ORDER_MORTON, wide=wide)[0], want_k), ("knn", k, wide)
    return b


def test_triangle_queries_vs_brute(A, orc, scenes):
    rng = np.random.default_rng(3)
    tris = np.concatenate([scenes.terrain(24, 5), random_tris(rng, 800, 12)])
    q = rng.uniform(-16, 30, (1500, 3)).astype(np.float32)
    check_queries(A, orc, 0, tris, q, 2.0)


def test_point_queries_vs_brute(A, orc, scenes):
    pts = scenes.point_cloud(6000, 2)
    pts[100:140] = pts[7]  # exact ties broken by index
    q = np.random.default_rng(4).uniform(-4, 68, (1200, 3)).astype(np.float32)
    check_queries(A, orc, 1, pts, q, 3.0)


def test_gaussian_queries_vs_brute(A, orc, scenes):
    g = scenes.gaussians(4000, 3)
    g[:, 3:7] *= np.float32(2.5)  # unnormalised quaternions
    q = np.random.default_rng(5).uniform(-4, 68, (800, 3)).astype(np.float32)
    check_queries(A, orc, 2, g, q, 1.5, ks=(1, 8, 32))


def _c2_pass(A, b, ref, q, what):
    """closest, kNN16 and radius 2.0 on binary and on 4-wide nodes, Morton order."""
    want_c, want_k = ref.knn(q, 0, THREADS), ref.knn(q, 16, THREADS)
    wo, wk = ref.radius(q, 2.0, THREADS)
    for wide in (False, True):
        log(f"C2 {what}: wide={wide}")
        assert eq(b.closest(q, order=A.ORDER_MORTON, wide=wide)[0], want_c), (what, "closest", wide)
        assert eq(b.knn(q, 16, order=A.ORDER_MORTON, wide=wide)[0], want_k), (what, "knn16", wide)
        o, k, _, _ = b.radius(q, 2.0, order=A.ORDER_MORTON, wide=wide)
        assert eq(o, wo) and eq(k, wk), (what, "radius", wide)


@pytest.mark.parametrize("builder", [0, 1])
def test_c2_full_size(A, orc, scenes, builder):
    """The whole 2M-triangle terrain workload: every result set of the query pass on
    binary and 4-wide nodes, then again after moving 1% of the grid vertices by
    N(0, 0.05) and a dirty refit, against the CPU BVH of the moved scene."""
    tris, q = scenes.workload_c2(seed=1)
    log(f"C2: 2M triangles, builder {builder}")
    b = A.Bvh(0, tris, builder)
    b.build_wide()
    _c2_pass(A, b, orc.CpuBvh(0, tris), q, "as built")
    moved, dirty = scenes.displace_vertices(1024, tris, 0.01, 0.05, seed=4)
    t = b.refit(moved, dirty)
    assert t.updated == len(np.unique(dirty))
    assert b.validate(moved) == ""
    _c2_pass(A, b, orc.CpuBvh(0, moved), q, "after the dirty refit")


def test_golden_fixtures(A):
    """GPU results equal the reference CPU code's stored outputs."""
    g = np.load(os.path.join(GOLD, "tri_mixed.npz"), allow_pickle=False)
    b = A.Bvh(0, g["tris"])
    b.build_wide()
    for wide in (False, True):
        assert eq(b.raycast(g["rays_down"], wide=wide)[0], g["ray_down"])
        assert eq(b.raycast(g["rays_random"], wide=wide, order=A.ORDER_MORTON)[0], g["ray_random"])
        assert eq(b.closest(g["points"], wide=wide)[0], g["closest"])
        assert eq(b.knn(g["points"], 6, wide=wide)[0], g["knn6"])
        o, k, _, _ = b.radius(g["points"], 3.0, wide=wide)
        assert eq(o, g["radius3_offsets"]) and eq(k, g["radius3_keys"])
    p = np.load(os.path.join(GOLD, "points.npz"), allow_pickle=False)
    b = A.Bvh(1, p["cloud"])
    assert eq(b.closest(p["points"])[0], p["closest"])
    assert eq(b.knn(p["points"], 8)[0], p["knn8"])
    o, k, _, _ = b.radius(p["points"], 2.5)
    assert eq(o, p["radius_offsets"]) and eq(k, p["radius_keys"])
    gg = np.load(os.path.join(GOLD, "gaussians.npz"), allow_pickle=False)
    b = A.Bvh(2, gg["gaussians"])
    assert eq(b.closest(gg["points"])[0], gg["closest"])
    assert eq(b.knn(gg["points"], 8)[0], gg["knn8"])
    o, k, _, _ = b.radius(gg["points"], 1.0)
    assert eq(o, gg["radius_offsets"]) and eq(k, gg["radius_keys"])


# ---- refit ------------------------------------------------------------------------------

def test_refit_dirty_rounds(A, orc, scenes):
    cells = 128
    tris = scenes.terrain(cells, 1)
    b = A.Bvh(0, tris)
    b.build_wide()
    q = scenes.query_points(np.array([-4, -4, -12], np.float32), np.array([132, 132, 12], np.float32), 3000, 9)
    cur = tris
    for rnd in range(4):
        log(f"refit round {rnd}")
        moved, dirty = scenes.displace_vertices(cells, cur, 0.01 * (rnd + 1), 0.05, seed=20 + rnd)
        t = b.refit(moved, dirty)
        assert t.updated == len(np.unique(dirty))
        assert b.validate(moved) == ""
        cur = moved
        cn_boxes_ok = b.validate(cur) == ""
        assert cn_boxes_ok
        assert eq(b.closest(q, wide=True)[0], orc.knn_brute(0, cur, q, 0, threads=THREADS))
    # a full refit after dirty rounds restores the original boxes exactly
    b.refit(tris)
    cn, _, _ = orc.lbvh_nodes(0, tris)
    assert eq(b.read()[0], cn)
    # dirty refit with duplicate ids and an empty list
    b.refit(tris, np.array([5, 5, 5], np.int32))
    b.refit(tris, np.array([], np.int32))
    assert eq(b.read()[0], cn)


@pytest.mark.parametrize("dirty_count", [40, 3000])
def test_refit_sparse_and_dense_paths(A, scenes, dirty_count):
    """Dirty refit takes the marked-path merge below 1 dirty leaf in 256 and the
    full bottom-up pass above it; both must give the full refit's node bytes."""
    tris = scenes.terrain(128, 5)  # 32,768 triangles
    rng = np.random.default_rng(dirty_count)
    idx = rng.choice(tris.shape[0], dirty_count, replace=False).astype(np.int32)
    moved = tris.copy()
    moved[idx] += rng.normal(0, 0.3, (dirty_count, 9)).astype(np.float32)
    a, b = A.Bvh(0, tris), A.Bvh(0, tris)
    a.refit(moved, idx)
    b.refit(moved)
    assert a.validate(moved) == ""
    assert eq(a.read()[0], b.read()[0])


def test_refit_points_and_gaussians(A, orc, scenes):
    rng = np.random.default_rng(8)
    for kind, items in ((1, scenes.point_cloud(5000, 4)), (2, scenes.gaussians(3000, 4))):
        b = A.Bvh(kind, items)
        moved = items.copy()
        idx = rng.choice(items.shape[0], 200, replace=False)
        moved[idx, 0:3] += rng.normal(0, 0.5, (200, 3)).astype(np.float32)
        b.refit(moved, idx)
        assert b.validate(moved) == ""
        q = rng.uniform(0, 64, (500, 3)).astype(np.float32)
        assert eq(b.knn(q, 8)[0], orc.knn_brute(kind, moved, q, 8, threads=THREADS))


# ---- error contract -----------------------------------------------------------------------

def test_errors(A, scenes):
    tris = scenes.terrain(8, 1)
    b = A.Bvh(0, tris)
    q = np.zeros((4, 3), np.float32)
    with pytest.raises(A.QuarryError):
        b.knn(q, 0)
    with pytest.raises(A.QuarryError):
        b.knn(q, 33)
    with pytest.raises(A.QuarryError):
        b.closest(q, wide=True)  # wide nodes not built
    with pytest.raises(A.QuarryError):
        A.Bvh(0, tris[:1])
    with pytest.raises(A.QuarryError):
        b.refit(tris[:-1])
    o, k, _, _ = b.radius(q, -1.0)
    assert o.tolist() == [0] * 5 and k.size == 0
    # empty query sets are fine
    assert b.closest(np.zeros((0, 3), np.float32))[0].shape == (0, 1, 2)


def test_radius_segments_above_workgroup_capacity(A, orc):
    """Segments longer than 8,192 keys take the device radix-sort path; a few
    queries see every one of 20,000 points (with ties on distance)."""
    rng = np.random.default_rng(21)
    pts = rng.uniform(-1, 1, (20000, 3)).astype(np.float32)
    pts[5000:5100] = pts[17]  # equal distances, ordered by index
    q = np.concatenate([np.zeros((3, 3)), rng.uniform(-3, 3, (60, 3))]).astype(np.float32)
    b = A.Bvh(1, pts)
    b.build_wide()
    want = orc.radius_brute(1, pts, q, 4.0, threads=THREADS)
    assert int(np.diff(want[0]).max()) > 8192
    for wide in (False, True):
        o, k, _, _ = b.radius(q, 4.0, order=A.ORDER_MORTON, wide=wide)
        assert eq(o, want[0]) and eq(k, want[1]), wide


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_traversal_modes_agree(A, orc, scenes, kind):
    """One lane per query and four lanes per query give the same bits."""
    rng = np.random.default_rng(31 + kind)
    items = {0: lambda: np.concatenate([scenes.terrain(20, 3), random_tris(rng, 400, 10)]),
             1: lambda: scenes.point_cloud(5000, 4), 2: lambda: scenes.gaussians(3000, 5)}[kind]()
    q = rng.uniform(-4, 68 if kind else 24, (900, 3)).astype(np.float32)
    b = A.Bvh(kind, items)
    b.build_wide()
    want = {k: orc.knn_brute(kind, items, q, k, threads=THREADS) for k in (0, 1, 5, 16, 32)}
    try:
        for mode in (1, 2, 3, 4, 5, 6, 0):
            A.check(A.lib().qr_set_traversal_mode(mode))
            for k, w in want.items():
                got = b.closest(q, order=A.ORDER_MORTON, wide=True)[0] if k == 0 else \
                    b.knn(q, k, order=A.ORDER_MORTON, wide=True)[0]
                assert eq(got, w), (mode, k)
    finally:
        A.check(A.lib().qr_set_traversal_mode(0))


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_radius_modes_agree(A, orc, scenes, kind):
    rng = np.random.default_rng(41 + kind)
    items = {0: lambda: np.concatenate([scenes.terrain(20, 3), random_tris(rng, 400, 10)]),
             1: lambda: scenes.point_cloud(5000, 4), 2: lambda: scenes.gaussians(3000, 5)}[kind]()
    q = rng.uniform(-4, 68 if kind else 24, (700, 3)).astype(np.float32)
    b = A.Bvh(kind, items)
    b.build_wide()
    want = orc.radius_brute(kind, items, q, 2.5, threads=THREADS)
    try:
        for mode in (1, 2, 3, 6, 0):
            A.check(A.lib().qr_set_traversal_mode(mode))
            o, k, _, _ = b.radius(q, 2.5, order=A.ORDER_MORTON, wide=True)
            assert eq(o, want[0]) and eq(k, want[1]), mode
    finally:
        A.check(A.lib().qr_set_traversal_mode(0))


def test_radius_repeated_calls_on_one_query_set(A, orc, scenes):
    """Radius calls on one resident query set compact their hits into keys sized by
    the previous call's total before the host reads the new total: a larger radius
    (more hits than that size, so the compaction is redone at the exact size), a
    smaller one (fits), and one whose densest queries pass the 64-hit capture (the
    fill walk writes into the same keys) must each give the oracle's CSR."""
    rng = np.random.default_rng(77)
    cloud = scenes.point_cloud(20000, 6)
    q = rng.uniform(0.0, 64.0, (3000, 3)).astype(np.float32)
    b = A.Bvh(1, cloud)
    b.build_wide()
    pts = A.PointBuffer(q)
    try:
        for r in (1.0, 3.0, 1.0, 6.0, 2.0):
            log(f"radius {r} on the resident query set")
            _, offs, keys = pts.radius(b, r, keep=True)
            want = orc.radius_brute(1, cloud, q, r, threads=THREADS)
            assert eq(offs, want[0]) and eq(keys, want[1]), r
    finally:
        pts.close()


def test_radius_wave_walk_wide_frontier(A, orc):
    """The one-wave-per-query capture on walks whose frontier is the whole tree: a
    radius holding every point (2,048-hit capture overflows, the fill walk takes
    over) and one holding about 3,000; equal to the four-lane walk and the oracle."""
    rng = np.random.default_rng(51)
    pts = rng.uniform(0, 1, (50000, 3)).astype(np.float32)
    pts[100:400] = pts[7]  # duplicates: equal keys, a deeper tree
    b = A.Bvh(1, pts)
    b.build_wide()
    for r, nq in ((2.0, 40), (0.25, 1500)):
        q = rng.uniform(-0.2, 1.2, (nq, 3)).astype(np.float32)
        want = orc.radius_brute(1, pts, q, r, threads=THREADS)
        seg = np.diff(want[0])
        assert seg.max() > 2048 and (r < 1 or seg.min() == 50000), r
        try:
            for mode in (0, 2):
                A.check(A.lib().qr_set_traversal_mode(mode))
                o, k, _, st = b.radius(q, r, order=A.ORDER_MORTON, wide=True, stats=True)
                assert eq(o, want[0]) and eq(k, want[1]), (r, mode)
                assert st.stack_overflows == 0
        finally:
            A.check(A.lib().qr_set_traversal_mode(0))


@pytest.mark.parametrize("kind", [1, 2])
def test_wave_walk_on_identical_items(A, orc, scenes, kind):
    """Every item identical: all box bounds tie, the whole tree is walked, and the k
    best are the k smallest indices -- in every traversal shape."""
    base = scenes.point_cloud(1, 3) if kind == 1 else scenes.gaussians(1, 3)
    items = np.repeat(base, 3000, axis=0)
    rng = np.random.default_rng(61)
    q = rng.uniform(-2, 66, (300, 3)).astype(np.float32)
    b = A.Bvh(kind, items)
    b.build_wide()
    want = {k: orc.knn_brute(kind, items, q, k, threads=THREADS) for k in (0, 7, 32)}
    try:
        for mode in (3, 2, 4, 5, 6, 0):
            A.check(A.lib().qr_set_traversal_mode(mode))
            for k, w in want.items():
                got, _, st = b.closest(q, wide=True, stats=True) if k == 0 else \
                    b.knn(q, k, order=A.ORDER_MORTON, wide=True, stats=True)
                assert eq(got, w), (mode, k)
                assert st.stack_overflows == 0
    finally:
        A.check(A.lib().qr_set_traversal_mode(0))


def test_tuned_shape_on_large_sets(A, orc, scenes):
    """From 65,536 queries the automatic mode tunes the shape on the first call (the
    trials write the output); the first and later calls give the oracle's bits."""
    tris = scenes.terrain(64, 3)
    b = A.Bvh(0, tris)
    b.build_wide()
    q = np.random.default_rng(71).uniform(-4, 68, (70000, 3)).astype(np.float32)
    q[:, 2] = np.random.default_rng(72).uniform(-12, 12, 70000).astype(np.float32)
    for k in (0, 8):
        want = orc.knn_brute(0, tris, q, k, threads=THREADS)
        for call in range(2):
            got = b.closest(q, order=A.ORDER_MORTON, wide=True)[0] if k == 0 else \
                b.knn(q, k, order=A.ORDER_MORTON, wide=True)[0]
            assert eq(got, want), (k, call)


def test_ray_shapes_agree(A, orc, scenes):
    """One lane and one wave per ray (forced, and tuned in mode 4) give the
    oracle's hit flags and t bits; with the hit primitive saved, the same primitive
    (the smallest index among equal-t hits)."""
    import ctypes

    V, F, tris = scenes.icosphere(levels=4, radius=10.0, noise=0.1, seed=2)
    items = np.concatenate([tris.reshape(-1, 9), scenes.terrain(24, 5)]).astype(np.float32)
    rays = scenes.sphere_rays(70000, seed=8)
    b = A.Bvh(0, items)
    b.build_wide()
    want = orc.ray_brute(items, rays, threads=THREADS)
    L = A.lib()
    dr = A.DeviceArray(rays)
    res = A.DeviceArray.empty(rays.shape[0], np.uint64)
    hit = A.DeviceArray.empty(rays.shape[0], np.int32)
    prims = {}
    try:
        for mode in (1, 3, 4, 0):
            A.check(L.qr_set_traversal_mode(mode))
            got = b.raycast(rays, order=A.ORDER_MORTON, wide=True)[0]
            assert np.array_equal(got, want), mode
            ms = ctypes.c_double()
            A.check(L.qr_raycast_save_device(b.h, dr.ptr, rays.shape[0], res.ptr, hit.ptr, A.ORDER_INPUT, 1,
                                             ctypes.byref(ms)))
            assert np.array_equal(res.download(), want), mode
            prims[mode] = hit.download()
    finally:
        A.check(L.qr_set_traversal_mode(0))
    assert all(np.array_equal(p, prims[1]) for p in prims.values())
    assert (prims[1] >= 0).sum() == int((want >> np.uint64(32)).sum())
