"""FP64 CPU restatement of the differentiable mode: the checker for the adjoint kernels.

TEST INFRASTRUCTURE ONLY (imported by tests/ and nothing else).  The reference has
no differentiable path; what fixes the answer here is the mathematics:
  * the value functions with a fixed primitive per query (closest point on a
    triangle by Ericson's regions, src/core/geometry.h:198-259; point distance;
    solid-ellipsoid distance, src/core/gaussian.h:23-131, with the root solved to
    FP64 precision; ray-plane t of the hit triangle) -- used for central finite
    differences;
  * the analytic adjoints of the same functions in FP64 (the CPU adjoint).
Pinned by tests/test_gpu_autograd.py against central differences (h = 1e-3) and by
closed-form spot checks (d dist2 / dq = 2 (q - p); dt / d origin for an axis-aligned
plane = -n / (n . d)).
"""
from __future__ import annotations

import numpy as np

F8 = np.float64


def _dot(a, b):
    return np.einsum("ij,ij->i", a, b)


# ---- closest point on a triangle (vectorised Ericson, FP64) --------------------------------

def tri_closest(q, a, b, c):
    """Per row: (dist2, weights (n, 3) of a, b, c, region id as in the GPU code)."""
    q, a, b, c = (np.asarray(x, F8) for x in (q, a, b, c))
    n = q.shape[0]
    ab, ac, ap = b - a, c - a, q - a
    d1, d2 = _dot(ab, ap), _dot(ac, ap)
    bp = q - b
    d3, d4 = _dot(ab, bp), _dot(ac, bp)
    cp = q - c
    d5, d6 = _dot(ab, cp), _dot(ac, cp)
    vc = d1 * d4 - d3 * d2
    vb = d5 * d2 - d1 * d6
    va = d3 * d6 - d5 * d4
    w = np.zeros((n, 3))
    region = np.full(n, -1)
    done = np.zeros(n, bool)

    def take(mask, wa, wb, wc, rid):
        m = mask & ~done
        w[m, 0], w[m, 1], w[m, 2] = (np.broadcast_to(x, (n,))[m] for x in (wa, wb, wc))
        region[m] = rid
        done[m] = True

    with np.errstate(divide="ignore", invalid="ignore"):
        take((d1 <= 0) & (d2 <= 0), 1.0, 0.0, 0.0, 0)
        take((d3 >= 0) & (d4 <= d3), 0.0, 1.0, 0.0, 1)
        t = d1 / (d1 - d3)
        take((vc <= 0) & (d1 >= 0) & (d3 <= 0), 1 - t, t, 0.0, 2)
        take((d6 >= 0) & (d5 <= d6), 0.0, 0.0, 1.0, 3)
        t = d2 / (d2 - d6)
        take((vb <= 0) & (d2 >= 0) & (d6 <= 0), 1 - t, 0.0, t, 4)
        t = (d4 - d3) / ((d4 - d3) + (d5 - d6))
        take((va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0), 0.0, 1 - t, t, 5)
        denom = va + vb + vc
        b1, b2 = vb / denom, vc / denom
        take(denom > 1e-20, 1 - b1 - b2, b1, b2, 6)
    # degenerate: nearest vertex
    rest = ~done
    if rest.any():
        dv = np.stack([((q - v) ** 2).sum(1) for v in (a, b, c)], 1)
        k = np.argmin(dv, 1)
        for j in range(3):
            m = rest & (k == j)
            w[m] = 0.0
            w[m, j] = 1.0
            region[m] = 7 + j
    p = w[:, 0:1] * a + w[:, 1:2] * b + w[:, 2:3] * c
    return ((q - p) ** 2).sum(1), w, region


def tri_dist2_fixed(q, tris, prim):
    t = np.asarray(tris, F8).reshape(-1, 3, 3)[prim]
    return tri_closest(q, t[:, 0], t[:, 1], t[:, 2])[0]


def tri_adjoint(q, tris, prim, g):
    """(dL/dq (n, 3), dL/dtris (P, 9)) for L = sum g * dist2, prim >= 0 rows only."""
    q = np.asarray(q, F8)
    tris = np.asarray(tris, F8).reshape(-1, 3, 3)
    ok = prim >= 0
    gq = np.zeros_like(q)
    gt = np.zeros((tris.shape[0], 9))
    t = tris[prim[ok]]
    _, w, _ = tri_closest(q[ok], t[:, 0], t[:, 1], t[:, 2])
    p = np.einsum("nk,nkj->nj", w, t)
    d = 2.0 * g[ok, None] * (q[ok] - p)
    gq[ok] = d
    for k in range(3):
        np.add.at(gt[:, 3 * k:3 * k + 3], prim[ok], -w[:, k:k + 1] * d)
    return gq, gt


# ---- points ------------------------------------------------------------------------------

def point_adjoint(q, cloud, prim, g):
    q = np.asarray(q, F8)
    cloud = np.asarray(cloud, F8)
    ok = prim >= 0
    gq = np.zeros_like(q)
    gc = np.zeros_like(cloud)
    d = 2.0 * g[ok, None] * (q[ok] - cloud[prim[ok]])
    gq[ok] = d
    np.add.at(gc, prim[ok], -d)
    return gq, gc


# ---- solid ellipsoids ----------------------------------------------------------------------

def _axes(quat):
    qn = quat / np.linalg.norm(quat, axis=1, keepdims=True)
    x, y, z, w = qn.T
    u = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w)], 1)
    v = np.stack([2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w)], 1)
    ww = np.stack([2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)], 1)
    return u, v, ww


def gauss_closest(q, g):
    """(dist2, root t, local y, closest local p, inside) with the root to FP64 precision."""
    q, g = np.asarray(q, F8), np.asarray(g, F8)
    m, quat, s = g[:, 0:3], g[:, 3:7], g[:, 7:10]
    u, v, w = _axes(quat)
    d = q - m
    y = np.stack([_dot(d, u), _dot(d, v), _dot(d, w)], 1)
    inside = ((y / s) ** 2).sum(1) <= 1.0
    ss = s * s
    lo = np.zeros(q.shape[0])
    hi = 2.0 * s.max(1) * np.linalg.norm(y, axis=1) + 1.0
    for _ in range(200):  # bisection: f decreasing in t
        mid = 0.5 * (lo + hi)
        f = ((s * y / (mid[:, None] + ss)) ** 2).sum(1) - 1.0
        lo = np.where(f > 0, mid, lo)
        hi = np.where(f > 0, hi, mid)
    t = 0.5 * (lo + hi)
    p = ss * y / (t[:, None] + ss)
    dist2 = ((y - p) ** 2).sum(1)
    dist2[inside] = 0.0
    t[inside] = 0.0
    return dist2, t, y, p, inside


def gauss_dist2_fixed(q, gs, prim):
    return gauss_closest(q, np.asarray(gs, F8)[prim])[0]


def gauss_adjoint(q, gs, prim, g):
    """(dL/dq, dL/dgaussians (P, 10): mean 3, quaternion 4, scale 3)."""
    q = np.asarray(q, F8)
    gs = np.asarray(gs, F8)
    ok = prim >= 0
    gq = np.zeros_like(q)
    gg = np.zeros((gs.shape[0], 10))
    sel = gs[prim[ok]]
    _, t, y, p, inside = gauss_closest(q[ok], sel)
    e = y - p
    u, v, w = _axes(sel[:, 3:7])
    gw = 2.0 * g[ok] * (~inside)
    dq = gw[:, None] * (e[:, 0:1] * u + e[:, 1:2] * v + e[:, 2:3] * w)
    gq[ok] = dq
    np.add.at(gg[:, 0:3], prim[ok], -dq)
    s = sel[:, 7:10]
    np.add.at(gg[:, 7:10], prim[ok], -gw[:, None] * t[:, None] * p * p / s ** 3)
    # rotation: dL/d(axis k) = gw e_k (q - m), chained through the normalised quaternion
    d = q[ok] - sel[:, 0:3]
    gu, gv, gww = (gw[:, None] * e[:, k:k + 1] * d for k in range(3))
    quat = sel[:, 3:7]
    nrm = np.linalg.norm(quat, axis=1, keepdims=True)
    x, yq, z, wq = (quat / nrm).T

    def dd(c0, c1, c2):
        return np.stack([c0, c1, c2], 1)

    z0 = np.zeros_like(x)
    du = [dd(z0, 2 * yq, 2 * z), dd(-4 * yq, 2 * x, -2 * wq), dd(-4 * z, 2 * wq, 2 * x), dd(z0, 2 * z, -2 * yq)]
    dv = [dd(2 * yq, -4 * x, 2 * wq), dd(2 * x, z0, 2 * z), dd(-2 * wq, -4 * z, 2 * yq), dd(-2 * z, z0, 2 * x)]
    dw = [dd(2 * z, -2 * wq, -4 * x), dd(2 * wq, 2 * z, -4 * yq), dd(2 * x, 2 * yq, z0), dd(2 * yq, -2 * x, z0)]
    gn = np.stack([_dot(gu, du[j]) + _dot(gv, dv[j]) + _dot(gww, dw[j]) for j in range(4)], 1)
    qn = quat / nrm
    gq4 = (gn - qn * (qn * gn).sum(1, keepdims=True)) / nrm
    np.add.at(gg[:, 3:7], prim[ok], gq4)
    return gq, gg


# ---- ray hit t ------------------------------------------------------------------------------

def ray_t_fixed(o, d, tris, prim):
    """Ray-plane t of the given triangle (FP64), per ray."""
    t3 = np.asarray(tris, F8).reshape(-1, 3, 3)[prim]
    a, b, c = t3[:, 0], t3[:, 1], t3[:, 2]
    n = np.cross(b - a, c - a)
    return _dot(n, a - np.asarray(o, F8)) / _dot(n, np.asarray(d, F8))


def ray_adjoint(o, d, tris, prim, g):
    """(dL/do, dL/dd, dL/dtris (P, 9)) for L = sum g * t over rays with prim >= 0."""
    o, d = np.asarray(o, F8), np.asarray(d, F8)
    tris = np.asarray(tris, F8).reshape(-1, 3, 3)
    ok = prim >= 0
    go, gd = np.zeros_like(o), np.zeros_like(d)
    gt = np.zeros((tris.shape[0], 9))
    t3 = tris[prim[ok]]
    a, b, c = t3[:, 0], t3[:, 1], t3[:, 2]
    n = np.cross(b - a, c - a)
    nd = _dot(n, d[ok])
    t = _dot(n, a - o[ok]) / nd
    k = g[ok, None] * n / nd[:, None]
    go[ok] = -k
    gd[ok] = -t[:, None] * k
    hp = o[ok] + t[:, None] * d[ok]
    # barycentrics of the hit point
    v0, v1, v2 = b - a, c - a, hp - a
    d00, d01, d11 = _dot(v0, v0), _dot(v0, v1), _dot(v1, v1)
    d20, d21 = _dot(v2, v0), _dot(v2, v1)
    den = d00 * d11 - d01 * d01
    bu = (d11 * d20 - d01 * d21) / den
    bv = (d00 * d21 - d01 * d20) / den
    for j, wj in enumerate((1 - bu - bv, bu, bv)):
        np.add.at(gt[:, 3 * j:3 * j + 3], prim[ok], wj[:, None] * k)
    return go, gd, gt
