"""Known answers and closed forms that pin the oracle's leaf math.

Cases follow the reference self-test (src/cpu/selftest.cpp:28-52, 100-123,
355-451) and add closed forms chosen so that a dropped term, a swapped sign or
a transposed operand fails: axis-aligned slabs, planar distances, a float64
bisection of the ellipsoid root equation.
"""
import math

import numpy as np
import pytest

TRI = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)  # z = 0, normal +z


def ray(o, d, tmin=1e-6, tmax=np.inf):
    return np.array([*o, *d, tmin, tmax], np.float32)


def test_triangle_known_answers(orc):
    ok, t = orc.tri_intersect(ray((0.25, 0.25, 1), (0, 0, -1)), TRI)
    assert ok and abs(t - 1.0) < 1e-5
    assert not orc.tri_intersect(ray((2, 2, 1), (0, 0, -1)), TRI)[0]
    assert not orc.tri_intersect(ray((0, 0, 1), (1, 0, 0)), TRI)[0]       # parallel
    assert not orc.tri_intersect(ray((0.25, 0.25, 1), (0, 0, 1)), TRI)[0]  # behind
    ok, t = orc.tri_intersect(ray((0.25, 0.25, -1), (0, 0, 1)), TRI)
    assert ok and abs(t - 1.0) < 1e-5
    # t scales with 1/|dir| (dir is not normalised) and tMax clips
    ok, t = orc.tri_intersect(ray((0.25, 0.25, 3), (0, 0, -2)), TRI)
    assert ok and t == np.float32(1.5)
    assert not orc.tri_intersect(ray((0.25, 0.25, 3), (0, 0, -2), tmax=1.4), TRI)[0]


def test_closest_point_known_answers(orc):
    cases = [((0.25, 0.25, 1.0), 1.0, 6), ((-1, -1, 0), 2.0, 0), ((2, 0, 0), 1.0, 1),
             ((0.5, 0.5, 0), 0.0, 5), ((0.5, -1, 0), 1.0, 2), ((-1, 0.5, 0), 1.0, 4),
             ((0, 2, 0), 1.0, 3), ((1, 1, 0), 0.5, 5)]
    for p, want, region in cases:
        d, reg, _ = orc.tri_closest(np.array(p, np.float32), TRI)
        assert abs(float(d) - want) < 1e-6, p
        assert reg == region, (p, reg)


def test_closest_point_is_planar_distance_inside(orc):
    rng = np.random.default_rng(3)
    for _ in range(200):
        u, v = rng.uniform(0.05, 0.45, 2)
        h = rng.uniform(-3, 3)
        p = np.array([u, v, h], np.float32)
        d, reg, b = orc.tri_closest(p, TRI)
        assert reg == 6
        assert abs(float(d) - float(p[2]) ** 2) <= 1e-6 * max(1.0, float(p[2]) ** 2)
        assert abs(b[0] - u) < 1e-5 and abs(b[1] - v) < 1e-5


def test_degenerate_triangle_falls_back_to_vertices(orc):
    tri = np.array([0, 0, 0, 1, 1, 1, 2, 2, 2], np.float32)  # collinear
    d, reg, _ = orc.tri_closest(np.array([3, 3, 3.5], np.float32), tri)
    assert reg in (3, 9)
    assert abs(float(d) - (1 + 1 + 1.5 ** 2)) < 1e-5


def test_box_distance_known_answers(orc):
    box = np.array([0, 0, 0, 1, 1, 1], np.float32)
    assert orc.box_dist2(np.array([0.5, 0.5, 0.5], np.float32), box) == 0.0
    assert orc.box_dist2(np.array([2, 0.5, 0.5], np.float32), box) == 1.0
    assert orc.box_dist2(np.array([2, 3, 1], np.float32), box) == 5.0
    assert orc.box_dist2(np.array([-1, -2, -3], np.float32), box) == 14.0


def G(mean, quat, scale):
    return np.array([*mean, *quat, *scale], np.float32)


def test_gaussian_axis_and_sphere_cases(orc):
    g = G((0, 0, 0), (0, 0, 0, 1), (1, 2, 3))
    assert abs(orc.gauss_dist2([5, 0, 0], g)[0] - 16.0) < 1e-4
    assert abs(orc.gauss_dist2([0, 4, 0], g)[0] - 4.0) < 1e-4
    assert abs(orc.gauss_dist2([0, 0, 7], g)[0] - 16.0) < 1e-4
    assert orc.gauss_dist2([0.5, 0.5, 0.5], g)[0] == 0.0
    s = math.sqrt(0.5)
    r = G((0, 0, 0), (0, 0, s, s), (1, 2, 3))  # 90 degrees about z
    assert abs(orc.gauss_dist2([4, 0, 0], r)[0] - 4.0) < 1e-3
    assert abs(orc.gauss_dist2([0, 5, 0], r)[0] - 16.0) < 1e-3
    sphere = G((0, 0, 0), (0, 0, 0, 1), (2, 2, 2))
    for p in ([5, 0, 0], [0, -3, 0], [1, 1, 1], [0, 0, 0], [3, 4, 0]):
        ln = float(np.linalg.norm(p))
        want = (ln - 2) ** 2 if ln > 2 else 0.0
        assert abs(orc.gauss_dist2(p, sphere)[0] - want) < 1e-4
    # unnormalised quaternions are normalised at use
    assert orc.gauss_dist2([5, 0, 0], G((0, 0, 0), (0, 0, 0, 7), (1, 2, 3)))[0] == \
        orc.gauss_dist2([5, 0, 0], g)[0]


def test_gaussian_box_is_lower_bound(orc):
    for i in range(200):
        g = G((0.5 * i - 50, 0.25 * i, 0.1 * i), (0.1 * i, 0.2, -0.1, 1.0), (0.1 + 0.01 * i, 0.05, 0.4))
        p = np.array([math.sin(0.3 * i) * 60, math.cos(0.7 * i) * 10, 0.5 * i], np.float32)
        assert orc.box_dist2(p, orc.gauss_bounds(g)) <= orc.gauss_dist2(p, g)[0] + 1e-5


def test_gaussian_float64_bisection_witness(orc):
    """Independent FP64 root of sum (a y / (t + a^2))^2 = 1 (selftest.cpp:400-451)."""
    worst = 0.0
    for i in range(300):
        ang = 0.37 * i
        g = G((math.cos(ang) * 20, 1, math.sin(ang) * 20), (math.sin(0.11 * i), 0.3, math.cos(0.7 * i), 1.0),
              (0.2 + 0.02 * i, 0.05 + 0.01 * i, 0.3))
        p = np.array([math.cos(0.9 * i) * 45, math.sin(0.3 * i) * 5, math.sin(0.5 * i) * 45], np.float32)
        q = g[3:7].astype(np.float64)
        q /= np.linalg.norm(q)
        x, y, z, w = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        yl = R.T @ (p.astype(np.float64) - g[0:3].astype(np.float64))
        a = g[7:10].astype(np.float64)
        if np.sum((yl / a) ** 2) <= 1.0:
            want = 0.0
        else:
            lo, hi = 0.0, 2 * a.max() * np.linalg.norm(yl)
            for _ in range(200):
                mid = 0.5 * (lo + hi)
                if np.sum((a * yl / (mid + a * a)) ** 2) - 1 > 0:
                    lo = mid
                else:
                    hi = mid
            t = 0.5 * (lo + hi)
            xc = a * a * yl / (t + a * a)
            want = float(np.sum((yl - xc) ** 2))
        got = float(orc.gauss_dist2(p, g)[0])
        rel = abs(math.sqrt(got) - math.sqrt(want)) / (math.sqrt(want) + a.max())
        worst = max(worst, rel)
    assert worst < 1e-4, worst


def test_ray_box_padding(orc):
    box = np.array([0, 0, 0, 1, 1, 1], np.float32)
    r = ray((0.5, 0.5, 5), (0, 0, -1))
    hit, _ = orc.lib().orc_ray_box, None
    t = np.zeros(1, np.float32)
    assert orc.lib().orc_ray_box(r.ctypes.data_as(orc._f32p), box.ctypes.data_as(orc._f32p), np.float32(10), t.ctypes.data_as(orc._f32p))
    assert t[0] == 4.0
    # pruned when the current best is closer than the entry (beyond the 1e-6 pad)
    assert not orc.lib().orc_ray_box(r.ctypes.data_as(orc._f32p), box.ctypes.data_as(orc._f32p), np.float32(3.9), t.ctypes.data_as(orc._f32p))
    # flat axis: |d| < 1e-20 tests containment only
    r2 = ray((1.5, 0.5, 5), (0, 0, -1))
    assert not orc.lib().orc_ray_box(r2.ctypes.data_as(orc._f32p), box.ctypes.data_as(orc._f32p), np.float32(10), t.ctypes.data_as(orc._f32p))
