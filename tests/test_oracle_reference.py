"""Oracle vs the reference's own CPU code on fresh random inputs (bitwise).

Runs where oracle/_ref/libquarry_ref.so was built from /root/reference
(`make -C oracle ref`); skipped elsewhere, where tests/golden pins the oracle.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def R(orc):
    r = orc.ref()
    if r is None:
        pytest.skip("reference CPU build (oracle/_ref) not present")
    return r


def f32(a):
    return np.ascontiguousarray(a, np.float32)


def test_leaf_math_bitwise(orc, R):
    rng = np.random.default_rng(11)
    for _ in range(3000):
        tri = f32(rng.normal(0, 3, 9))
        p = f32(rng.normal(0, 5, 3))
        a = orc.tri_closest(p, tri)[0]
        b = R.ref_tri_closest(p.ctypes.data_as(orc._f32p), tri.ctypes.data_as(orc._f32p))
        assert np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32)
        r = f32([*rng.normal(0, 5, 3), *rng.normal(0, 1, 3), 1e-6, np.inf])
        t1 = np.zeros(1, np.float32)
        ok1 = R.ref_tri_intersect(r.ctypes.data_as(orc._f32p), tri.ctypes.data_as(orc._f32p), t1.ctypes.data_as(orc._f32p))
        ok2, t2 = orc.tri_intersect(r, tri)
        assert bool(ok1) == ok2
        if ok2:
            assert np.float32(t2).view(np.uint32) == t1.view(np.uint32)[0]
        g = f32([*rng.normal(0, 2, 3), *rng.normal(0, 1, 4), *np.exp(rng.uniform(-3, 0.5, 3))])
        a = orc.gauss_dist2(p, g)[0]
        b = R.ref_gauss_dist2(p.ctypes.data_as(orc._f32p), g.ctypes.data_as(orc._f32p))
        assert np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32)


def test_reference_terrain_scene(orc, R):
    cap = 2 * 64 * 64
    tris = np.zeros((cap, 9), np.float32)
    b6 = np.zeros(6, np.float32)
    n = R.ref_scene(0, 64, 0, 1, tris.ctypes.data_as(orc._f32p), cap, b6.ctypes.data_as(orc._f32p))
    rays = np.zeros((20000, 8), np.float32)
    R.ref_rays(b6.ctypes.data_as(orc._f32p), 1, 20000, 3, rays.ctypes.data_as(orc._f32p))
    pts = np.zeros((2000, 3), np.float32)
    R.ref_query_points(b6.ctypes.data_as(orc._f32p), 2000, 3, pts.ctypes.data_as(orc._f32p))
    h = R.ref_bvh_new(tris.ctypes.data_as(orc._f32p), n)
    want = np.zeros(20000, np.uint64)
    R.ref_bvh_ray(h, rays.ctypes.data_as(orc._f32p), 20000, want.ctypes.data_as(orc._u64p), 4)
    knn = np.zeros((2000, 16, 2), np.uint32)
    R.ref_bvh_knn(h, pts.ctypes.data_as(orc._f32p), 2000, 16, knn.ctypes.data_as(orc._u32p), 4)
    R.ref_bvh_free(h)
    bvh = orc.CpuBvh(orc.PRIM_TRI, tris[:n])
    assert np.array_equal(bvh.ray(rays, threads=4), want)
    assert np.array_equal(bvh.knn(pts, 16, threads=4), knn)
