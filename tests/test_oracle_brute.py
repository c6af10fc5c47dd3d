"""Tree engines equal brute force bitwise; brute force agrees with float64 numpy.

The float64 numpy brute force is an independent statement of the problem
(closest point on a triangle by projection/clamping onto the plane, edges and
vertices; point distances), so a sign or index mistake in the oracle's FP32
restatement shows up as a different argmin or a distance off by far more than
rounding.
"""
import numpy as np
import pytest


def closest_on_tri64(p, t):
    """Exact closest point on triangle t (3x3) to points p (N,3), float64."""
    a, b, c = t
    best = np.full(p.shape[0], np.inf)
    n = np.cross(b - a, c - a)
    nn = n @ n
    if nn > 1e-30:
        # interior projection
        d = (p - a) @ n / nn
        q = p - np.outer(d, n)
        v0, v1, v2 = b - a, c - a, q - a
        d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
        d20, d21 = v2 @ v0, v2 @ v1
        den = d00 * d11 - d01 * d01
        v = (d11 * d20 - d01 * d21) / den
        w = (d00 * d21 - d01 * d20) / den
        inside = (v >= 0) & (w >= 0) & (v + w <= 1)
        best = np.where(inside, np.sum((p - q) ** 2, axis=1), best)
    for s, e in ((a, b), (b, c), (c, a)):
        se = e - s
        tt = np.clip((p - s) @ se / max(se @ se, 1e-300), 0, 1)
        q = s + np.outer(tt, se)
        best = np.minimum(best, np.sum((p - q) ** 2, axis=1))
    return best


@pytest.fixture(scope="module")
def small_scene():
    rng = np.random.default_rng(21)
    a = rng.uniform(-10, 10, (300, 3))
    tris = np.concatenate([a, a + rng.uniform(-2, 2, (300, 3)), a + rng.uniform(-2, 2, (300, 3))], 1)
    pts = rng.uniform(-12, 12, (200, 3))
    return tris.astype(np.float32), pts.astype(np.float32)


def test_closest_matches_float64(orc, small_scene):
    tris, pts = small_scene
    got = orc.knn_brute(orc.PRIM_TRI, tris, pts, 0)
    d64 = np.stack([closest_on_tri64(pts.astype(np.float64), tris[i].reshape(3, 3).astype(np.float64))
                    for i in range(tris.shape[0])], 1)
    want_d = d64.min(1)
    gd = got[:, 0, 0].view(np.float32).astype(np.float64)
    assert np.all(np.abs(gd - want_d) <= 1e-4 * np.maximum(1.0, want_d))
    # argmin agrees except on float32 near-ties
    gi = got[:, 0, 1].astype(np.int64)
    mism = gi != d64.argmin(1)
    assert np.all(np.abs(d64[np.arange(len(gi)), gi] - want_d)[mism] <= 1e-4 * np.maximum(1.0, want_d[mism]))


def test_point_queries_match_float64(orc):
    rng = np.random.default_rng(4)
    cloud = rng.uniform(-5, 5, (2000, 3)).astype(np.float32)
    q = rng.uniform(-6, 6, (300, 3)).astype(np.float32)
    got = orc.knn_brute(orc.PRIM_POINT, cloud, q, 5)
    d = ((q[:, None, :] - cloud[None, :, :]) ** 2)  # float32, same op order per axis
    d32 = (d[..., 0] + d[..., 1]) + d[..., 2]
    order = np.lexsort((np.broadcast_to(np.arange(2000), d32.shape), d32), axis=1)[:, :5]
    assert np.array_equal(got[:, :, 1].astype(np.int64), order)
    assert np.array_equal(got[:, :, 0], np.take_along_axis(d32, order, 1).view(np.uint32))


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_tree_equals_brute(orc, kind):
    rng = np.random.default_rng(100 + kind)
    n = 4000
    if kind == 0:
        a = rng.uniform(-30, 30, (n, 3))
        items = np.concatenate([a, a + rng.normal(0, 1, (n, 3)), a + rng.normal(0, 1, (n, 3))], 1)
    elif kind == 1:
        items = rng.uniform(-30, 30, (n, 3))
        items[10:40] = items[5]
    else:
        items = np.concatenate([rng.uniform(-30, 30, (n, 3)), rng.normal(size=(n, 4)),
                                np.exp(rng.uniform(-3, 0, (n, 3)))], 1)
    items = items.astype(np.float32)
    q = rng.uniform(-33, 33, (300, 3)).astype(np.float32)
    bvh = orc.CpuBvh(kind, items)
    for k in (0, 1, 7, 32):
        assert np.array_equal(bvh.knn(q, k), orc.knn_brute(kind, items, q, k, threads=4)), k
    o1, k1 = bvh.radius(q, 2.5)
    o2, k2 = orc.radius_brute(kind, items, q, 2.5, threads=4)
    assert np.array_equal(o1, o2) and np.array_equal(k1, k2)
    if kind == 1:
        kd = orc.CpuKd(items)
        assert np.array_equal(kd.knn(q, 9), bvh.knn(q, 9))
        o3, k3 = kd.radius(q, 2.5)
        assert np.array_equal(o3, o1) and np.array_equal(k3, k1)


def test_radius_edge_cases(orc):
    items = np.zeros((3, 3), np.float32)
    q = np.zeros((2, 3), np.float32)
    o, k = orc.radius_brute(1, items, q, -1.0)
    assert o.tolist() == [0, 0, 0] and k.size == 0
    o, k = orc.radius_brute(1, items, q, 0.0)  # dist2 == r2 is inside
    assert o.tolist() == [0, 3, 6]
    assert [int(x) & 0xffffffff for x in k[:3]] == [0, 1, 2]
