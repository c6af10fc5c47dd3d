"""Acceptance suite of the Python module, in the order of the reference's own
script (python/test_quarry.py:170-688): one demo terrain (grid 32, 2,048
triangles) goes through queries, the SAH builder, rays, vertex and partial
updates, topology edits (delete + append, move-only, mid and edge inserts), then a
20,000-point cloud and a 2,000-Gaussian cloud go through the same sequence, and
finally every bad input of the reference's list must raise ValueError.

Each test carries the reference check names it covers.  Where the reference
compares with an FP64 brute force under a tolerance (python/test_quarry.py:18-157),
this suite keeps that criterion and adds the stronger one of this
implementation: the results equal the oracle's FP32 brute force bit for bit.

The tests share one state (module fixture) and run in file order, as the script
does: the scene the update tests edit is the scene the query tests built.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from oracle import adjoint as J

pytestmark = pytest.mark.gpu
THREADS = 8


@pytest.fixture(scope="module")
def st():
    import quarry

    return {"quarry": quarry, "rng": np.random.default_rng(7)}


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def same(a, b):
    return np.array_equal(_bits(a), _bits(b))


def tri_items(tris):
    return np.ascontiguousarray(tris, np.float32).reshape(-1, 9)


def fp64_tri_matrix(points, tris):
    """(len(points), len(tris)) FP64 squared distances, Ericson's regions."""
    items = tri_items(tris).astype(np.float64)
    n, m = len(points), len(items)
    q = np.repeat(np.asarray(points, np.float64), m, axis=0)
    prim = np.tile(np.arange(m), n)
    return J.tri_dist2_fixed(q, items, prim).reshape(n, m)


def fp64_gauss_matrix(points, means, rot, scale):
    g = np.concatenate([means, rot, scale], 1).astype(np.float64)
    n, m = len(points), len(g)
    q = np.repeat(np.asarray(points, np.float64), m, axis=0)
    prim = np.tile(np.arange(m), n)
    return J.gauss_dist2_fixed(q, g, prim).reshape(n, m)


def expect_closest(kind, items, points, d2, idx):
    exp = O.knn_brute(kind, items, points, 0, threads=THREADS)
    assert np.array_equal(_bits(d2), exp[:, 0, 0]), "closest dist2 bits"
    assert np.array_equal(idx, exp[:, 0, 1].view(np.int32)), "closest ids"


def expect_knn(kind, items, points, k, d2, idx):
    exp = O.knn_brute(kind, items, points, k, threads=THREADS)
    assert np.array_equal(_bits(d2), exp[:, :, 0]), "knn dist2 bits"
    assert np.array_equal(idx, exp[:, :, 1].view(np.int32)), "knn ids"


def expect_radius(kind, items, points, r, offs, d2, idx):
    eo, ek = O.radius_brute(kind, items, points, r, threads=THREADS)
    assert np.array_equal(offs, eo.astype(np.int64)), "radius offsets"
    assert np.array_equal(_bits(d2), (ek >> np.uint64(32)).astype(np.uint32)), "radius dist2 bits"
    assert np.array_equal(idx, (ek & np.uint64(0xFFFFFFFF)).astype(np.int32)), "radius ids"


def check_closest_like_reference(scene, points, tris):
    """The reference's criterion (dist2 within 2e-4 of FP64, ids minimal up to
    ties) and this implementation's (bit-equal to the FP32 brute force)."""
    D = fp64_tri_matrix(points, tris)
    ref_d2 = D.min(axis=1)
    d2, tri = scene.closest(points)
    assert np.allclose(d2, ref_d2, rtol=2e-4, atol=1e-6)
    assert np.all(D[np.arange(len(points)), tri] <= ref_d2 * (1 + 2e-4) + 1e-6)
    expect_closest(0, tri_items(tris), points, d2, tri)
    return d2, tri, D


def rays_of(origins, dirs):
    r = np.zeros((len(origins), 8), np.float32)
    r[:, 0:3] = origins
    r[:, 3:6] = dirs
    r[:, 6] = np.float32(1e-6)
    r[:, 7] = np.inf
    return r


def expect_raycast(tris, origins, dirs, hit, t):
    exp = O.ray_brute(tri_items(tris), rays_of(origins, dirs), threads=THREADS)
    eh = (exp >> np.uint64(32)) != 0
    assert np.array_equal(hit, eh), "ray hit flags"
    assert np.array_equal(_bits(t[hit]), (exp[eh] & np.uint64(0xFFFFFFFF)).astype(np.uint32)), "ray t bits"
    assert np.all(np.isinf(t[~hit])), "miss = inf"


def assert_same_closest(a, b, points):
    da, ia = a.closest(points)
    db, ib = b.closest(points)
    assert np.array_equal(ia, ib) and same(da, db)


def assert_same_knn(a, b, points, k):
    da, ia = a.knn(points, k)
    db, ib = b.knn(points, k)
    assert np.array_equal(ia, ib) and same(da, db)


def assert_same_raycast(a, b, origins, dirs):
    ha, ta = a.raycast(origins, dirs)
    hb, tb = b.raycast(origins, dirs)
    assert np.array_equal(ha, hb) and same(ta, tb)


# ---- triangle scene -------------------------------------------------------------------------

def test_scene_build(st):
    """GPU available, demo_mesh shapes, triangle_count, node_count, validate, build_ms,
    bounds order."""
    quarry = st["quarry"]
    assert quarry.describe()
    vertices, indices = quarry.demo_mesh("terrain", grid=32)
    assert vertices.shape == (2048 * 3, 3) and indices.shape == (2048, 3)
    scene = quarry.Scene(vertices, indices)
    assert scene.triangle_count == 2048
    assert scene.node_count == 2 * 2048 - 1
    assert scene.validate() == ""
    assert scene.build_ms > 0.0
    lo, hi = np.array(scene.bounds["min"]), np.array(scene.bounds["max"])
    assert np.all(lo <= hi)
    tris = vertices[indices].astype(np.float32)
    # the bounds are the union of the triangles' boxes
    assert np.array_equal(_bits(lo), _bits(tris.reshape(-1, 3).min(axis=0)))
    assert np.array_equal(_bits(hi), _bits(tris.reshape(-1, 3).max(axis=0)))
    points = (lo + (hi - lo) * st["rng"].random((300, 3))).astype(np.float32)
    st.update(vertices=vertices, indices=indices, scene=scene, lo=lo, hi=hi, tris=tris, points=points)


def test_scene_queries(st):
    """closest distance, closest ids (ties allowed), last_query_ms, knn shapes, knn first =
    closest, knn dist2 sorted, knn dist2 values, radius offsets, radius dist2 sorted per
    segment, radius bound, radius sets."""
    scene, points, tris = st["scene"], st["points"], st["tris"]
    dist2, tri, D = check_closest_like_reference(scene, points, tris)
    assert scene.last_query_ms > 0.0
    k = 4
    kd, kid = scene.knn(points, k)
    assert kd.shape == (len(points), k) and kid.shape == (len(points), k)
    assert np.array_equal(kid[:, 0], tri)
    assert np.all(np.diff(kd, axis=1) >= 0)
    assert np.allclose(kd, np.sort(D, axis=1)[:, :k], rtol=2e-4, atol=1e-6)
    expect_knn(0, tri_items(tris), points, k, kd, kid)
    radius = 0.3 * float(np.max(st["hi"] - st["lo"]))
    offsets, rd2, rid = scene.radius(points, radius)
    assert offsets.shape == (len(points) + 1,) and offsets[0] == 0 and offsets[-1] == len(rid)
    for i in range(len(points)):
        seg = slice(offsets[i], offsets[i + 1])
        assert np.all(np.diff(rd2[seg]) >= 0)
        # the reference's FP64 set, away from the boundary band of FP32 rounding
        inside = set(np.nonzero(D[i] <= radius * radius * (1 - 1e-4))[0].tolist())
        outside = set(np.nonzero(D[i] > radius * radius * (1 + 1e-4))[0].tolist())
        got = set(rid[seg].tolist())
        assert inside <= got and not (got & outside)
    assert np.all(rd2 <= radius * radius + 1e-5)
    expect_radius(0, tri_items(tris), points, radius, offsets, rd2, rid)
    st.update(dist2=dist2, tri=tri, kd=kd, kid=kid, k=k, radius=radius, offsets=offsets, rd2=rd2, rid=rid)


def test_morton_order_bitwise(st):
    """closest morton bitwise, knn morton bitwise, radius morton bitwise."""
    scene, points = st["scene"], st["points"]
    d2, tri = scene.closest(points, order="morton")
    assert np.array_equal(tri, st["tri"]) and same(d2, st["dist2"])
    kd, kid = scene.knn(points, st["k"], order="morton")
    assert np.array_equal(kid, st["kid"]) and same(kd, st["kd"])
    off, rd2, rid = scene.radius(points, st["radius"], order="morton")
    assert np.array_equal(off, st["offsets"]) and np.array_equal(rid, st["rid"]) and same(rd2, st["rd2"])


def test_sah_builder_bitwise(st):
    """sah validate, sah node_count, sah closest / knn / radius bitwise, sah build_stats
    has bin breakdown."""
    quarry, points = st["quarry"], st["points"]
    sah = quarry.Scene(st["vertices"], st["indices"], builder="sah")
    assert sah.validate() == ""
    assert sah.node_count == 2 * 2048 - 1
    d2, tri = sah.closest(points)
    assert np.array_equal(tri, st["tri"]) and same(d2, st["dist2"])
    kd, kid = sah.knn(points, st["k"])
    assert np.array_equal(kid, st["kid"]) and same(kd, st["kd"])
    off, rd2, rid = sah.radius(points, st["radius"])
    assert np.array_equal(off, st["offsets"]) and np.array_equal(rid, st["rid"]) and same(rd2, st["rd2"])
    assert "sah_bin_ms" in sah.build_stats


def test_raycast(st):
    """raycast hits, raycast t, raycast miss = inf, triangles constructor."""
    quarry, scene, lo, hi = st["quarry"], st["scene"], st["lo"], st["hi"]
    origins = (lo + (hi - lo) * st["rng"].random((200, 3))).astype(np.float32)
    origins[:, 1] = hi[1] + 1.0
    dirs = np.tile(np.array([0.0, -1.0, 0.0], np.float32), (200, 1))
    hit, t = scene.raycast(origins, dirs)
    assert hit.sum() > 150
    expect_raycast(st["tris"], origins, dirs, hit, t)
    face = quarry.Scene(st["tris"])
    assert face.triangle_count == 2048
    st.update(origins=origins, dirs=dirs, face_scene=face)


def test_update_vertices(st):
    """update keeps triangle count, update validate, update refit_ms, update uploaded
    everything, update bounds, update closest."""
    scene, vertices, indices = st["scene"], st["vertices"], st["indices"]
    lifted = vertices.copy()
    lifted[:, 1] += 0.1
    scene.update_vertices(lifted)
    assert scene.triangle_count == 2048
    assert scene.validate() == ""
    assert scene.refit_ms > 0.0
    assert scene.refit_stats["updated"] == 2048
    assert np.allclose(np.array(scene.bounds["max"]), lifted.max(axis=0), atol=1e-5)
    check_closest_like_reference(scene, st["points"], lifted[indices].astype(np.float32))
    st["lifted"] = lifted


def test_partial_updates(st):
    """partial update validate, partial updated count, partial update closest, second
    partial update validate / closest (the second edit shares leaves and ancestors with
    the first)."""
    scene, indices = st["scene"], st["indices"]
    partial = st["lifted"][indices].astype(np.float32)
    dirty = np.array([0, 5, 7, 2047], np.int32)
    partial[dirty] += np.array([0.0, 3.0, 0.0], np.float32)
    scene.update_triangles(partial, dirty=dirty)
    assert scene.validate() == ""
    assert scene.refit_stats["updated"] == len(dirty)
    check_closest_like_reference(scene, st["points"], partial)
    again = partial.copy()
    dirty2 = np.array([0, 5, 7, 2047, 100], np.int32)
    again[dirty2] += np.array([0.0, 3.0, 0.0], np.float32)
    scene.update_triangles(again, dirty=dirty2)
    assert scene.validate() == ""
    assert scene.refit_stats["updated"] == len(dirty2)
    check_closest_like_reference(scene, st["points"], again)
    st.update(partial=partial, again=again)


def test_topology_delete_append(st):
    """topo triangle count, topo validate, topo edit counts, topo closest / knn / raycast
    bitwise against a fresh build."""
    quarry, scene, lo, hi = st["quarry"], st["scene"], st["lo"], st["hi"]
    again = st["again"]
    deleted = np.array([3, 100, 700], np.int32)
    keep = np.ones(len(again), bool)
    keep[deleted] = False
    extra = (lo + 0.2 * (hi - lo) * st["rng"].random((150, 3, 3))).astype(np.float32)
    new_tris = np.concatenate([again[keep], extra], axis=0)
    topo = scene.update_topology(new_tris, deleted=deleted)
    assert scene.triangle_count == len(new_tris)
    assert scene.validate() == ""
    assert topo["deleted"] == 3 and topo["appended"] == 150 and topo["moved"] == 0
    fresh = quarry.Scene(new_tris)
    assert_same_closest(scene, fresh, st["points"])
    assert_same_knn(scene, fresh, st["points"], 8)
    assert_same_raycast(scene, fresh, st["origins"], st["dirs"])
    expect_knn(0, tri_items(new_tris), st["points"], 8, *scene.knn(st["points"], 8))
    st["new_tris"] = new_tris


def test_topology_move_only(st):
    """topo move-only uses refit, topo move validate, topo move closest bitwise."""
    quarry, scene = st["quarry"], st["scene"]
    shifted = st["new_tris"].copy()
    moved = np.array([0, 5, 11], np.int32)
    shifted[moved] += np.array([0.0, 0.5, 0.0], np.float32)
    m = scene.update_topology(shifted, moved=moved)
    assert not m["full_rebuild"] and m["refit_ms"] > 0.0
    assert scene.validate() == ""
    assert_same_closest(scene, quarry.Scene(shifted), st["points"])
    st["shifted"] = shifted


def test_topology_mid_and_edge_insert(st):
    """mid-insert validate, mid-insert counts, mid-insert closest / knn / raycast bitwise,
    edge-insert validate, edge-insert closest bitwise, stale mesh binding refused."""
    quarry, scene, lo, hi, rng = st["quarry"], st["scene"], st["lo"], st["hi"], st["rng"]
    shifted = st["shifted"]
    n_mid, base = 60, len(shifted)
    pos = np.sort(rng.choice(base + n_mid, size=n_mid, replace=False)).astype(np.int32)
    extra = (lo + 0.2 * (hi - lo) * rng.random((n_mid, 3, 3))).astype(np.float32)
    mask = np.zeros(base + n_mid, bool)
    mask[pos] = True
    mid = np.empty((base + n_mid, 3, 3), np.float32)
    mid[mask] = extra
    mid[~mask] = shifted
    t = scene.update_topology(mid, inserted=pos)
    assert scene.validate() == ""
    assert t["inserted"] == n_mid and t["appended"] == 0
    fresh = quarry.Scene(mid)
    assert_same_closest(scene, fresh, st["points"])
    assert_same_knn(scene, fresh, st["points"], 8)
    assert_same_raycast(scene, fresh, st["origins"], st["dirs"])
    expect_raycast(mid, st["origins"], st["dirs"], *scene.raycast(st["origins"], st["dirs"]))
    # one insert before everything, one after everything.  np.insert takes positions
    # in the old array; update_topology takes them in the new scene (docs/api.md:226-235),
    # where the two new triangles sit at 0 and len(edge) - 1.  (The reference script
    # passes the np.insert positions [0, len(mid)]: that declares the old last triangle
    # inserted and the new tail triangle a survivor, whose geometry the tree keeps.)
    edge_extra = (lo + 0.2 * (hi - lo) * rng.random((2, 3, 3))).astype(np.float32)
    edge = np.insert(mid, np.array([0, len(mid)]), edge_extra, axis=0)
    edge_pos = np.array([0, len(edge) - 1], np.int32)
    assert same(edge[edge_pos], edge_extra)
    scene.update_topology(edge, inserted=edge_pos)
    assert scene.validate() == ""
    assert_same_closest(scene, quarry.Scene(np.ascontiguousarray(edge, np.float32)), st["points"])
    expect_closest(0, tri_items(edge), st["points"], *scene.closest(st["points"]))
    # the vertex binding no longer describes the scene
    with pytest.raises(ValueError):
        scene.update_vertices(st["lifted"])
    st.update(mid_tris=mid, mid_extra=extra)


# ---- point cloud ----------------------------------------------------------------------------

def test_point_cloud_build_and_queries(st):
    """point_count, point node_count, point validate, point build_ms, point bounds, point
    closest dist2 bitwise, point closest ids, point last_query_ms, point knn bitwise,
    point radius offsets, point radius bitwise, point closest morton bitwise."""
    quarry, lo, hi, rng = st["quarry"], st["lo"], st["hi"], st["rng"]
    npts, nq = 20000, 200
    pts = (lo + (hi - lo) * rng.random((npts, 3))).astype(np.float32)
    cloud = quarry.PointCloud(pts)
    assert cloud.point_count == npts
    assert cloud.node_count == 2 * npts - 1
    assert cloud.validate() == ""
    assert cloud.build_ms > 0.0
    plo, phi = np.array(cloud.bounds["min"]), np.array(cloud.bounds["max"])
    assert np.array_equal(_bits(plo), _bits(pts.min(axis=0))) and np.array_equal(_bits(phi), _bits(pts.max(axis=0)))
    q = (lo + (hi - lo) * rng.random((nq, 3))).astype(np.float32)
    diff = pts[None, :, :] - q[:, None, :]
    # FP32 squared distance summed x, y, z in order: the GPU's own operation order
    d2 = (diff * diff).sum(axis=-1, dtype=np.float32)
    rows = np.arange(nq)
    best = d2.argmin(axis=1)
    cd2, cidx = cloud.closest(q)
    assert same(cd2, d2[rows, best])
    assert np.array_equal(cidx, best.astype(np.int32))
    assert cloud.last_query_ms > 0.0
    ck = 4
    kd, kid = cloud.knn(q, ck)
    for i in range(nq):
        order = np.argsort(d2[i], kind="stable")[:ck]  # (dist2, index) order
        assert np.array_equal(kid[i], order.astype(np.int32)) and same(kd[i], d2[i][order])
    r = 0.08 * float(np.linalg.norm(phi - plo))
    off, rd2, rid = cloud.radius(q, r)
    assert off.shape == (nq + 1,) and off[0] == 0 and off[-1] == len(rid)
    for i in range(nq):
        want = np.nonzero(d2[i] <= np.float32(r) * np.float32(r))[0]
        want = want[np.argsort(d2[i][want], kind="stable")]
        assert np.array_equal(rid[off[i]:off[i + 1]], want.astype(np.int32))
        assert same(rd2[off[i]:off[i + 1]], d2[i][want])
    expect_radius(1, pts, q, r, off, rd2, rid)
    md2, midx = cloud.closest(q, order="morton")
    assert np.array_equal(midx, cidx) and same(md2, cd2)
    st.update(cloud=cloud, pts=pts, q=q)


def test_point_cloud_updates(st):
    """point update validate, point update refit_ms, point update uploaded everything,
    point update closest bitwise, point partial validate, point partial updated count,
    point partial closest bitwise."""
    cloud, pts, q = st["cloud"], st["pts"], st["q"]
    moved = pts.copy()
    moved[:, 1] += np.float32(0.1)
    cloud.update_points(moved)
    assert cloud.validate() == ""
    assert cloud.refit_ms > 0.0
    assert cloud.refit_stats["updated"] == len(pts)
    expect_closest(1, moved, q, *cloud.closest(q))
    partial = moved.copy()
    dirty = np.array([0, 3, 17, len(pts) - 1], np.int32)
    partial[dirty] += np.array([0.0, 0.5, 0.0], np.float32)
    cloud.update_points(partial, dirty=dirty)
    assert cloud.validate() == ""
    assert cloud.refit_stats["updated"] == len(dirty)
    expect_closest(1, partial, q, *cloud.closest(q))
    st["partial_pts"] = partial


def test_point_cloud_topology(st):
    """point topo count, point topo validate, point topo edit counts, point topo closest
    bitwise, point mid-insert validate, point mid-insert counts, point mid-insert closest
    bitwise."""
    quarry, cloud, q, lo, hi, rng = st["quarry"], st["cloud"], st["q"], st["lo"], st["hi"], st["rng"]
    partial = st["partial_pts"]
    pdel = np.array([0, 7, 10, 555], np.int32)
    keep = np.ones(len(partial), bool)
    keep[pdel] = False
    new_pts = np.concatenate([partial[keep], (lo + (hi - lo) * rng.random((500, 3))).astype(np.float32)])
    t = cloud.update_topology(new_pts, deleted=pdel)
    assert cloud.point_count == len(new_pts)
    assert cloud.validate() == ""
    assert t["deleted"] == 4 and t["appended"] == 500
    assert_same_closest(cloud, quarry.PointCloud(new_pts), q)
    n_ins, base = 200, len(new_pts)
    pos = np.sort(rng.choice(base + n_ins, size=n_ins, replace=False)).astype(np.int32)
    mask = np.zeros(base + n_ins, bool)
    mask[pos] = True
    mid = np.empty((base + n_ins, 3), np.float32)
    mid[mask] = (lo + (hi - lo) * rng.random((n_ins, 3))).astype(np.float32)
    mid[~mask] = new_pts
    t = cloud.update_topology(mid, inserted=pos)
    assert cloud.validate() == ""
    assert t["inserted"] == n_ins and t["appended"] == 0
    assert_same_closest(cloud, quarry.PointCloud(mid), q)
    expect_closest(1, mid, q, *cloud.closest(q))


# ---- Gaussian cloud -------------------------------------------------------------------------

def test_gaussian_cloud_build_and_queries(st):
    """gaussian_count, gaussian node_count, gaussian validate, gaussian build_ms, gaussian
    closest dist2, gaussian closest ids (ties allowed), gaussian last_query_ms, gaussian
    closest morton bitwise, gaussian knn shapes, gaussian knn dist2, gaussian radius
    offsets, gaussian radius sets."""
    quarry, rng = st["quarry"], st["rng"]
    ngs, ngq = 2000, 200
    means, rot, scale = quarry.demo_gaussians("sheet", ngs, seed=3)
    g = quarry.GaussianCloud(means, rot, scale)
    assert g.gaussian_count == ngs
    assert g.node_count == 2 * ngs - 1
    assert g.validate() == ""
    assert g.build_ms > 0.0
    blo, bhi = np.array(g.bounds["min"]), np.array(g.bounds["max"])
    gq = (blo + (bhi - blo) * rng.random((ngq, 3))).astype(np.float32)
    items = np.concatenate([means, rot, scale], 1).astype(np.float32)
    D = fp64_gauss_matrix(gq, means, rot, scale)
    rows = np.arange(ngq)
    ref = D.min(axis=1)
    d2, idx = g.closest(gq)
    assert np.allclose(d2, ref, rtol=1e-4, atol=1e-6)
    assert np.all(D[rows, idx] <= ref * (1 + 1e-4) + 1e-6)
    expect_closest(2, items, gq, d2, idx)
    assert g.last_query_ms > 0.0
    md2, midx = g.closest(gq, order="morton")
    assert np.array_equal(midx, idx) and same(md2, d2)
    gk = 4
    kd, kid = g.knn(gq, gk)
    assert kd.shape == (ngq, gk) and kid.shape == (ngq, gk)
    assert np.allclose(kd, np.sort(D, axis=1)[:, :gk], rtol=1e-4, atol=1e-6)
    expect_knn(2, items, gq, gk, kd, kid)
    r = 0.03 * float(np.linalg.norm(bhi - blo))
    off, rd2, rid = g.radius(gq, r)
    assert off.shape == (ngq + 1,) and off[0] == 0 and off[-1] == len(rid)
    r2, band = r * r, 1e-3 * r * r
    for i in range(ngq):
        seg = slice(off[i], off[i + 1])
        got = set(rid[seg].tolist())
        assert set(np.nonzero(D[i] <= r2 - band)[0].tolist()) <= got
        assert not (got & set(np.nonzero(D[i] > r2 + band)[0].tolist()))
        assert np.all(np.diff(rd2[seg]) >= 0)
    expect_radius(2, items, gq, r, off, rd2, rid)
    st.update(gcloud=g, gmeans=means, grot=rot, gscale=scale, gq=gq)


def test_gaussian_cloud_updates(st):
    """gaussian update validate, gaussian update uploaded everything, gaussian update
    closest, gaussian partial validate, gaussian partial updated count, gaussian partial
    closest."""
    g, means, rot, scale, gq = st["gcloud"], st["gmeans"], st["grot"], st["gscale"], st["gq"]
    moved = means.copy()
    moved[:, 1] += np.float32(0.1)
    g.update(moved, rot, scale)
    assert g.validate() == ""
    assert g.refit_stats["updated"] == len(means)
    D = fp64_gauss_matrix(gq, moved, rot, scale)
    rows = np.arange(len(gq))
    d2, idx = g.closest(gq)
    assert np.allclose(d2, D.min(axis=1), rtol=1e-4, atol=1e-6)
    assert np.all(D[rows, idx] <= D.min(axis=1) * (1 + 1e-4) + 1e-6)
    expect_closest(2, np.concatenate([moved, rot, scale], 1), gq, d2, idx)
    partial = moved.copy()
    dirty = np.array([0, 7, len(means) - 1], np.int32)
    partial[dirty] += np.array([0.0, 0.5, 0.0], np.float32)
    g.update(partial, rot, scale, dirty=dirty)
    assert g.validate() == ""
    assert g.refit_stats["updated"] == len(dirty)
    expect_closest(2, np.concatenate([partial, rot, scale], 1), gq, *g.closest(gq))
    st["gpartial"] = partial


def test_gaussian_cloud_topology(st):
    """gaussian topo count, gaussian topo validate, gaussian topo edit counts, gaussian
    topo closest dist2 bitwise, gaussian topo ids are minimizers, gaussian mid-insert
    validate, gaussian mid-insert counts, gaussian mid-insert dist2 bitwise."""
    quarry, g, rng, gq = st["quarry"], st["gcloud"], st["rng"], st["gq"]
    means, rot, scale, partial = st["gmeans"], st["grot"], st["gscale"], st["gpartial"]
    gdel = np.array([1, 2, 100], np.int32)
    keep = np.ones(len(means), bool)
    keep[gdel] = False
    kmin, kmax = means[keep].min(axis=0), means[keep].max(axis=0)
    add_m = (kmin + (kmax - kmin) * rng.random((300, 3))).astype(np.float32)
    add_r = rot[rng.integers(0, 10, 300)]
    add_s = scale[rng.integers(0, 10, 300)] * np.float32(0.5)
    new_m = np.concatenate([partial[keep], add_m])
    new_r = np.concatenate([rot[keep], add_r])
    new_s = np.concatenate([scale[keep], add_s])
    t = g.update_topology(new_m, new_r, new_s, deleted=gdel)
    assert g.gaussian_count == len(new_m)
    assert g.validate() == ""
    assert t["deleted"] == 3 and t["appended"] == 300
    fresh = quarry.GaussianCloud(new_m, new_r, new_s)
    td2, tid = g.closest(gq)
    fd2, fid = fresh.closest(gq)
    assert same(td2, fd2) and np.array_equal(tid, fid)
    D = fp64_gauss_matrix(gq, new_m, new_r, new_s)
    assert np.all(D[np.arange(len(gq)), tid] <= td2 * (1 + 1e-4) + 1e-6)
    n_ins, base = 40, len(new_m)
    pos = np.sort(rng.choice(base + n_ins, size=n_ins, replace=False)).astype(np.int32)
    mask = np.zeros(base + n_ins, bool)
    mask[pos] = True
    lo, hi = means.min(axis=0), means.max(axis=0)
    cols = []
    for ins, old, w in (((lo + (hi - lo) * rng.random((n_ins, 3))).astype(np.float32), new_m, 3),
                        (rot[rng.integers(0, 10, n_ins)], new_r, 4),
                        (scale[rng.integers(0, 10, n_ins)] * np.float32(0.5), new_s, 3)):
        a = np.empty((base + n_ins, w), np.float32)
        a[mask] = ins
        a[~mask] = old
        cols.append(a)
    t = g.update_topology(*cols, inserted=pos)
    assert g.validate() == ""
    assert t["inserted"] == n_ins and t["appended"] == 0
    md2, _ = g.closest(gq)
    nd2, _ = quarry.GaussianCloud(*cols).closest(gq)
    assert same(md2, nd2)
    expect_closest(2, np.concatenate(cols, 1), gq, *g.closest(gq))
    st.update(gmoved_bad=means[:10])


# ---- bad inputs -----------------------------------------------------------------------------

def test_bad_inputs_raise_value_error(st):
    """bad inputs raise ValueError (python/test_quarry.py:664-683)."""
    quarry = st["quarry"]
    scene, cloud, g = st["scene"], st["cloud"], st["gcloud"]
    points, q, pts = st["points"], st["q"], st["pts"]
    mid, mid_extra, new_tris = st["mid_tris"], st["mid_extra"], st["new_tris"]
    means, rot, scale = st["gmeans"], st["grot"], st["gscale"]
    dirs = st["dirs"]
    bad_scene = quarry.Scene(mid)
    bad_big = np.concatenate([mid, mid_extra[:2]], axis=0)
    calls = (
        lambda: quarry.Scene(np.zeros((5, 3), np.float32)),
        lambda: quarry.PointCloud(np.zeros((1, 3), np.float32)),
        lambda: scene.knn(points, 0),
        lambda: scene.knn(points, 33),
        lambda: cloud.knn(q, 33),
        lambda: scene.raycast(points, dirs[:-1]),
        lambda: scene.update_triangles(st["partial"][:10]),
        lambda: cloud.update_points(pts[:10]),
        lambda: g.update(means[:10], rot, scale),
        lambda: scene.update_topology(new_tris[:-1]),
        lambda: bad_scene.update_topology(bad_big, inserted=np.array([5, 3], np.int32)),
        lambda: bad_scene.update_topology(bad_big, inserted=np.array([0, len(bad_big)], np.int32)),
        lambda: scene.update_topology(new_tris, deleted=np.array([5, 3], np.int32)),
        lambda: scene.update_topology(new_tris, deleted=np.array([0, 10 ** 6], np.int32)),
        lambda: cloud.update_topology(pts[:10]),
        lambda: quarry.GaussianCloud(means, rot, scale[:10]),
        lambda: quarry.GaussianCloud(means, rot, -scale),
        lambda: st["face_scene"].update_vertices(st["lifted"]),
    )
    for i, call in enumerate(calls):
        print("bad input", i, flush=True)
        with pytest.raises(ValueError):
            call()
    # every object is still usable after the refused calls
    assert scene.validate() == "" and cloud.validate() == "" and g.validate() == ""
