"""Benchmark scene generators: shapes, determinism, documented parameters."""
import numpy as np


def test_terrain_shape_and_determinism(scenes):
    a = scenes.terrain(16, 3)
    b = scenes.terrain(16, 3)
    assert a.shape == (2 * 16 * 16, 9) and a.dtype == np.float32
    assert np.array_equal(a, b)
    assert not np.array_equal(a, scenes.terrain(16, 4))
    t = a.reshape(-1, 3, 3)
    assert t[..., 0].min() == 0 and t[..., 0].max() == 16
    assert np.abs(t[..., 2]).max() <= 8.25


def test_displacement_marks_incident_triangles(scenes):
    tris = scenes.terrain(32, 1)
    moved, dirty = scenes.displace_vertices(32, tris, 0.05, 0.05, seed=2)
    changed = np.nonzero(np.any(moved != tris, axis=1))[0]
    assert np.array_equal(np.sort(dirty), changed)


def test_other_generators(scenes):
    c = scenes.city(500, 1)
    assert c.shape == (2 * 128 * 128 + 10 * 500, 9)
    g = scenes.gaussians(100, 1)
    assert g.shape == (100, 10) and np.all(g[:, 7:] > 0)
    assert np.allclose(np.linalg.norm(g[:, 3:7], axis=1), 1, atol=1e-5)
    p = scenes.point_cloud(1000, 1)
    assert p.shape == (1000, 3)
    v, f, t = scenes.icosphere(3)
    assert f.shape == (20 * 4 ** 3, 3) and t.shape == (20 * 4 ** 3, 9)
    r = scenes.sphere_rays(100)
    assert np.allclose(np.linalg.norm(r[:, 3:6], axis=1), 1, atol=1e-5)
