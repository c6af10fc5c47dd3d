"""The FP64 CPU adjoint (oracle/adjoint.py) pinned by mathematics, on the CPU.

Central finite differences (h = 1e-3) of the fixed-primitive value functions,
closed forms, and agreement of the FP64 closest-point regions with the C oracle's
FP32 ones.  A plausible slip (a dropped weight, a sign, a transposed axis, a
missing normalisation term) fails one of these.
"""
import numpy as np
import pytest

from oracle import adjoint as J

H = 1e-3


def central(fun, x, idx, h):
    xp, xm = x.copy(), x.copy()
    xp.flat[idx] += h
    xm.flat[idx] -= h
    return (fun(xp).sum() - fun(xm).sum()) / (2 * h)


def fd(fun, x, idx, h=H):
    """FP64 central difference at step h, Richardson-refined with h/2 (removes the
    h^2 truncation term, which at h = 1e-3 alone can reach 1e-4 of a small
    component where the function curves strongly)."""
    return (4.0 * central(fun, x, idx, h / 2) - central(fun, x, idx, h)) / 3.0


def rel_err(a, b, floor):
    """|a - b| relative to |b|, with components far below the array's typical
    magnitude judged against `floor` instead (callers pass 5% of its rms)."""
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def rms(x):
    return float(np.sqrt(np.mean(np.square(x))))


def _mesh(rng, n=40):
    a = rng.uniform(-5, 5, (n, 3))
    return np.concatenate([a, a + rng.normal(0, 1.5, (n, 3)), a + rng.normal(0, 1.5, (n, 3))], 1)


def test_tri_regions_cover_all_and_match_c_oracle(orc):
    rng = np.random.default_rng(1)
    tris = _mesh(rng, 300)
    q = rng.uniform(-8, 8, (300, 3))
    d2, w, region = J.tri_closest(q, tris[:, 0:3], tris[:, 3:6], tris[:, 6:9])
    assert np.all(region >= 0)
    assert np.allclose(w.sum(1), 1.0)
    assert np.all(w >= -1e-12)
    agree = 0
    for i in range(300):
        rd, rreg, _ = orc.tri_closest(q[i].astype(np.float32), tris[i].astype(np.float32))
        assert abs(float(rd) - d2[i]) <= 1e-4 * (1 + d2[i])
        agree += int(rreg) == int(region[i])
    assert agree >= 297  # FP32 vs FP64 may flip a region only at a boundary


def test_tri_adjoint_vs_fd():
    rng = np.random.default_rng(2)
    tris = _mesh(rng, 30)
    q = rng.uniform(-8, 8, (200, 3))
    prim = rng.integers(0, 30, 200)
    g = rng.uniform(0.5, 2.0, 200)
    gq, gt = J.tri_adjoint(q, tris, prim, g)
    f_q = lambda x: g * J.tri_dist2_fixed(x, tris, prim)
    f_t = lambda x: g * J.tri_dist2_fixed(q, x, prim)
    for idx in rng.choice(q.size, 60, replace=False):
        assert rel_err(gq.flat[idx], fd(f_q, q, idx), 1e-3) < 1e-5
    for idx in rng.choice(tris.size, 60, replace=False):
        assert rel_err(gt.flat[idx], fd(f_t, tris, idx), 0.05 * rms(gt)) < 1e-4
    # a grazing ray (|cos| ~ 2e-3): the adjoint still equals a fine difference
    a = np.array([[0.0, 0, 0, 4, 0, 0, 0, 4, 0]])
    o1, d1 = np.array([[1.0, 1.0, -0.01]]), np.array([[1.0, 0.0, 0.002]])
    _, _, g1 = J.ray_adjoint(o1, d1, a, np.array([0]), np.array([1.0]))
    for idx in range(9):
        assert rel_err(g1.flat[idx], central(lambda x: J.ray_t_fixed(o1, d1, x, np.array([0])), a, idx, 1e-7), 1e-3) < 1e-4


def test_point_adjoint_closed_form():
    rng = np.random.default_rng(3)
    cloud = rng.normal(size=(50, 3))
    q = rng.normal(size=(80, 3))
    prim = rng.integers(0, 50, 80)
    g = np.ones(80)
    gq, gc = J.point_adjoint(q, cloud, prim, g)
    assert np.allclose(gq, 2 * (q - cloud[prim]))
    assert np.allclose(gc.sum(0), -gq.sum(0))


def test_gauss_value_known_answers():
    # sphere of radius 2 at the origin: distance from (5,0,0) is 3, inside is 0
    g = np.array([[0, 0, 0, 0, 0, 0, 1, 2, 2, 2]], float)
    d2, t, *_ = J.gauss_closest(np.array([[5.0, 0, 0]]), g)
    assert abs(d2[0] - 9.0) < 1e-12 and abs(t[0] - 2 * 3.0) < 1e-9  # t = s (|y| - s)
    assert J.gauss_closest(np.array([[1.0, 0.5, 0]]), g)[0][0] == 0.0
    # axis-aligned ellipsoid, point on the long axis: distance |y| - a
    g2 = np.array([[1, 2, 3, 0, 0, 0, 1, 4, 1, 0.5]], float)
    assert abs(J.gauss_closest(np.array([[1 + 7.0, 2, 3]]), g2)[0][0] - 9.0) < 1e-10


def test_gauss_adjoint_vs_fd():
    rng = np.random.default_rng(4)
    n = 40
    gs = np.concatenate([rng.uniform(-5, 5, (n, 3)), rng.normal(size=(n, 4)) * 1.7,
                         np.exp(rng.uniform(np.log(0.3), np.log(2.0), (n, 3)))], 1)
    q = rng.uniform(-9, 9, (150, 3))
    prim = rng.integers(0, n, 150)
    g = rng.uniform(0.5, 2.0, 150)
    gq, gg = J.gauss_adjoint(q, gs, prim, g)
    f_q = lambda x: g * J.gauss_dist2_fixed(x, gs, prim)
    f_g = lambda x: g * J.gauss_dist2_fixed(q, x, prim)
    for idx in rng.choice(q.size, 60, replace=False):
        assert rel_err(gq.flat[idx], fd(f_q, q, idx), 1e-3) < 1e-5
    for idx in rng.choice(gs.size, 120, replace=False):
        assert rel_err(gg.flat[idx], fd(f_g, gs, idx), 0.05 * rms(gg)) < 1e-4, idx % 10


def test_ray_adjoint_closed_form_and_fd():
    # plane z = 2 spanned by one big triangle; ray from the origin straight up
    tris = np.array([[-10, -10, 2, 10, -10, 2, 0, 10, 2]], float)
    o = np.array([[0.5, 0.5, 0.0]])
    d = np.array([[0.0, 0.0, 1.0]])
    go, gd, gt = J.ray_adjoint(o, d, tris, np.array([0]), np.array([1.0]))
    assert np.allclose(go, [[0, 0, -1]]) and np.allclose(gd, [[0, 0, -2]])
    assert np.allclose(gt.reshape(3, 3)[:, 2].sum(), 1.0)  # lifting the plane by h lifts t by h
    rng = np.random.default_rng(5)
    tris = _mesh(rng, 20)
    prim = rng.integers(0, 20, 100)
    t3 = tris.reshape(-1, 3, 3)[prim]
    target = t3.mean(1)
    o = target + rng.normal(0, 6, (100, 3))
    d = target - o
    # central differences at h = 1e-3 lose accuracy as 1/cos^3 at grazing incidence
    # (checked separately below); the sampled rays meet their plane at |cos| > 0.2
    nrm = np.cross(t3[:, 1] - t3[:, 0], t3[:, 2] - t3[:, 0])
    cos = np.abs(np.einsum("ij,ij->i", nrm, d)) / np.linalg.norm(nrm, axis=1) / np.linalg.norm(d, axis=1)
    keep = cos > 0.2
    o, d, prim = o[keep], d[keep], prim[keep]
    g = rng.uniform(0.5, 2.0, o.shape[0])
    go, gd, gt = J.ray_adjoint(o, d, tris, prim, g)
    f_o = lambda x: g * J.ray_t_fixed(x, d, tris, prim)
    f_d = lambda x: g * J.ray_t_fixed(o, x, tris, prim)
    f_t = lambda x: g * J.ray_t_fixed(o, d, x, prim)
    for idx in rng.choice(o.size, 40, replace=False):
        assert rel_err(go.flat[idx], fd(f_o, o, idx), 1e-3) < 1e-5
        assert rel_err(gd.flat[idx], fd(f_d, d, idx), 1e-3) < 1e-4
    for idx in rng.choice(tris.size, 60, replace=False):
        assert rel_err(gt.flat[idx], fd(f_t, tris, idx), 0.05 * rms(gt)) < 1e-4
    # a grazing ray (|cos| ~ 2e-3): the adjoint still equals a fine difference
    a = np.array([[0.0, 0, 0, 4, 0, 0, 0, 4, 0]])
    o1, d1 = np.array([[1.0, 1.0, -0.01]]), np.array([[1.0, 0.0, 0.002]])
    _, _, g1 = J.ray_adjoint(o1, d1, a, np.array([0]), np.array([1.0]))
    for idx in range(9):
        assert rel_err(g1.flat[idx], central(lambda x: J.ray_t_fixed(o1, d1, x, np.array([0])), a, idx, 1e-7), 1e-3) < 1e-4
