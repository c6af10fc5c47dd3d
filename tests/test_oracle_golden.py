"""The oracle reproduces the reference CPU code's outputs bit for bit.

tests/golden/*.npz hold inputs and the outputs of the reference's own CPU path
(src/cpu/bvh.cpp, point_kd.cpp, gauss_bvh.cpp compiled from its sources by
tests/golden/make_golden.py).  Every oracle engine (brute force, median-split
BVH, kd-tree) must match them exactly.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_triangles_rays(orc):
    g = load("tri_mixed.npz")
    tris = g["tris"]
    bvh = orc.CpuBvh(orc.PRIM_TRI, tris)
    for key, out in (("rays_down", "ray_down"), ("rays_random", "ray_random")):
        assert np.array_equal(bvh.ray(g[key]), g[out])
        assert np.array_equal(orc.ray_brute(tris, g[key], threads=4), g[out])
    assert int(np.count_nonzero(g["ray_down"] >> np.uint64(32))) > 1000


def test_triangles_queries(orc):
    g = load("tri_mixed.npz")
    tris, pts = g["tris"], g["points"]
    bvh = orc.CpuBvh(orc.PRIM_TRI, tris)
    assert np.array_equal(bvh.knn(pts, 0), g["closest"])
    assert np.array_equal(orc.knn_brute(orc.PRIM_TRI, tris, pts, 0, threads=4), g["closest"])
    assert np.array_equal(bvh.knn(pts, 6), g["knn6"])
    assert np.array_equal(orc.knn_brute(orc.PRIM_TRI, tris, pts, 6, threads=4), g["knn6"])
    for o, k in (bvh.radius(pts, 3.0), orc.radius_brute(orc.PRIM_TRI, tris, pts, 3.0, threads=4)):
        assert np.array_equal(o, g["radius3_offsets"])
        assert np.array_equal(k, g["radius3_keys"])


def test_points(orc):
    g = load("points.npz")
    cloud, pts = g["cloud"], g["points"]
    kd = orc.CpuKd(cloud)
    bvh = orc.CpuBvh(orc.PRIM_POINT, cloud)
    for eng in (kd, bvh):
        assert np.array_equal(eng.knn(pts, 0), g["closest"])
        assert np.array_equal(eng.knn(pts, 8), g["knn8"])
        o, k = eng.radius(pts, 2.5)
        assert np.array_equal(o, g["radius_offsets"]) and np.array_equal(k, g["radius_keys"])
    assert np.array_equal(orc.knn_brute(orc.PRIM_POINT, cloud, pts, 8, threads=4), g["knn8"])


def test_gaussians(orc):
    g = load("gaussians.npz")
    gs, pts = g["gaussians"], g["points"]
    bvh = orc.CpuBvh(orc.PRIM_GAUSS, gs)
    assert np.array_equal(bvh.knn(pts, 0), g["closest"])
    assert np.array_equal(bvh.knn(pts, 8), g["knn8"])
    o, k = bvh.radius(pts, 1.0)
    assert np.array_equal(o, g["radius_offsets"]) and np.array_equal(k, g["radius_keys"])
    d = np.array([orc.gauss_dist2(pts[i % 300], gs[i])[0] for i in range(gs.shape[0])], np.float32)
    assert np.array_equal(d.view(np.uint32), g["pair_dist2"].view(np.uint32))
