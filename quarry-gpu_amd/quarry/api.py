"""The public Python surface of `quarry`: Scene, PointCloud, GaussianCloud.

Mirrors the reference's pybind module (bindings/python/quarry_py.cpp:887-1050,
docs/api.md:277-392): the same classes, constructors, properties, methods,
argument names, defaults, return shapes and error types --
  * invalid arguments (shapes, k outside [1, 32], negative radius, unknown order or
    builder, invalid topology edits) raise ValueError;
  * device failures raise RuntimeError (quarry.QuarryError).
Every call goes through the C ABI of libquarry_hip.so; nothing here computes a
query result on the CPU.

Additions over the reference (its Python surface is binary-only,
docs/design.md:122): `wide=` on every query selects the 4-wide nodes (built on
first use), `order=` on `raycast`, and `build_wide()`.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import _abi as A

_ORDERS = {"input": A.ORDER_INPUT, "morton": A.ORDER_MORTON}
_RAY_ORDERS = {"input": A.ORDER_INPUT, "morton": A.ORDER_MORTON,
               "morton_direction": A.ORDER_MORTON_DIRECTION}


def _order(s: str, table=_ORDERS) -> int:
    if s not in table:
        raise ValueError("order must be " + " or ".join(repr(k) for k in table))
    return table[s]


def _builder(s: str) -> int:
    # parseSahBuilder (quarry_py.cpp:42-46)
    if s == "lbvh":
        return A.BUILDER_LBVH
    if s == "sah":
        return A.BUILDER_SAH
    raise ValueError("builder must be 'lbvh' or 'sah'")


def _f32(a, what: str) -> np.ndarray:
    try:
        return np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    except (TypeError, ValueError) as e:
        raise ValueError(f"{what} must be a float32 array") from e


def _points(a, what: str = "points") -> np.ndarray:
    p = _f32(a, what)
    if p.ndim != 2 or p.shape[1] != 3:
        raise ValueError(f"{what} must have shape (N, 3)")
    return p


def _triangles(a) -> np.ndarray:
    t = _f32(a, "triangles")
    if t.ndim != 3 or t.shape[1:] != (3, 3):
        raise ValueError("triangles must have shape (M, 3, 3)")
    if t.shape[0] < 2:
        raise ValueError("at least 2 triangles are required")
    return t


def _mesh(vertices, indices) -> np.ndarray:
    v = _f32(vertices, "vertices")
    if v.ndim != 2 or v.shape[1] != 3:
        raise ValueError("vertices must have shape (N, 3)")
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.int32))
    if idx.ndim != 2 or idx.shape[1] != 3:
        raise ValueError("indices must have shape (M, 3)")
    if idx.shape[0] < 2:
        raise ValueError("at least 2 triangles are required")
    if idx.size and (idx.min() < 0 or idx.max() >= v.shape[0]):
        raise ValueError("indices out of range")
    return np.ascontiguousarray(v[idx])


def _gaussians(means, rot, scale) -> np.ndarray:
    m, q, s = _f32(means, "means"), _f32(rot, "rot"), _f32(scale, "scale")
    for arr in (m, s):
        if arr.ndim != 2 or arr.shape[1] != 3:
            raise ValueError("means and scales must have shape (N, 3)")
    if q.ndim != 2 or q.shape[1] != 4:
        raise ValueError("rotations must have shape (N, 4) as xyzw quaternions")
    if not (q.shape[0] == m.shape[0] == s.shape[0]):
        raise ValueError("means, rotations and scales must have the same length")
    if not np.all(s > 0):
        raise ValueError("semi-axes must be positive")
    return np.ascontiguousarray(np.concatenate([m, q, s], axis=1))


def _ids(a) -> np.ndarray | None:
    if a is None:
        return None
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))


def _root_box(bvh: A.Bvh) -> dict:
    """Scene bounds as the reference reports them (the union of the primitive
    boxes): the root node's box, read from the device."""
    node = np.zeros(8, np.float32)
    A.check(A.lib().qr_dev_download(A.ptr(node), bvh.info().d_nodes, 32))
    return {"min": tuple(float(x) for x in node[0:3]), "max": tuple(float(x) for x in node[3:6])}


def _bits_to_f32(u: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(u, dtype=np.uint32).view(np.float32)


class _Index:
    """Shared machinery of Scene / PointCloud / GaussianCloud."""

    _kind = A.PRIM_TRI

    def _build(self, items: np.ndarray, builder: str) -> None:
        self._builder = _builder(builder)
        self._items = items
        t0 = time.perf_counter()
        self._bvh = A.Bvh(self._kind, items, self._builder)
        self._wall_build_ms = 1000.0 * (time.perf_counter() - t0)
        self._timings = self._bvh.timings
        self._refit = None
        self._refit_ms = 0.0
        self._last_query_ms = 0.0
        self._wide = False

    # -- properties (quarry_py.cpp:921-939) --
    @property
    def node_count(self) -> int:
        return self._bvh.info().node_count

    @property
    def depth(self) -> int:
        return self._bvh.info().depth

    @property
    def build_ms(self) -> float:
        return float(self._timings.total_ms)

    @property
    def last_query_ms(self) -> float:
        return self._last_query_ms

    @property
    def refit_ms(self) -> float:
        return self._refit_ms

    @property
    def refit_stats(self) -> dict:
        r = self._refit
        keys = ("upload_ms", "scatter_ms", "tree_ms", "mark_ms", "merge_ms", "total_ms")
        if r is None:
            return {**{k: 0.0 for k in keys}, "updated": 0}
        return {"upload_ms": r.upload_ms, "scatter_ms": r.gather_ms, "tree_ms": r.tree_ms,
                "mark_ms": r.mark_ms, "merge_ms": r.merge_ms, "total_ms": r.total_ms,
                "updated": int(r.updated)}

    @property
    def build_stats(self) -> dict:
        t = self._timings
        return {"morton_ms": t.morton_ms, "sort_ms": t.sort_ms, "scatter_ms": t.scatter_ms,
                "tree_ms": t.tree_ms, "refit_ms": t.refit_ms, "total_ms": t.total_ms,
                "sah_bin_ms": t.sah_bin_ms, "sah_split_ms": t.sah_split_ms,
                "sah_partition_ms": t.sah_part_ms}

    @property
    def bounds(self) -> dict:
        return _root_box(self._bvh)

    def validate(self) -> str:
        return self._bvh.validate(self._items)

    def build_wide(self) -> float:
        """Collapses the tree into 4-wide nodes; returns the collapse time (ms)."""
        ms = self._bvh.build_wide()
        self._wide = True
        return ms

    def _use_wide(self, wide: bool) -> bool:
        if wide and not self._wide:
            self.build_wide()
        return bool(wide)

    # -- queries (quarry_py.cpp:250-360) --
    def closest(self, points, order: str = "input", wide: bool = False):
        p = _points(points)
        o = _order(order)
        rec, ms, _ = self._bvh.closest(p, o, self._use_wide(wide))
        self._last_query_ms = ms
        return _bits_to_f32(rec[:, 0, 0]).copy(), rec[:, 0, 1].view(np.int32).copy()

    def knn(self, points, k: int, order: str = "input", wide: bool = False):
        if not (1 <= int(k) <= 32):
            raise ValueError("k must be in [1, 32]")
        p = _points(points)
        o = _order(order)
        rec, ms, _ = self._bvh.knn(p, int(k), o, self._use_wide(wide))
        self._last_query_ms = ms
        return _bits_to_f32(rec[:, :, 0]).copy(), rec[:, :, 1].view(np.int32).copy()

    def radius(self, points, radius: float, order: str = "input", wide: bool = False):
        if not (float(radius) >= 0.0):
            raise ValueError("radius must be >= 0")
        p = _points(points)
        o = _order(order)
        offs, keys, ms, _ = self._bvh.radius(p, float(radius), o, self._use_wide(wide))
        self._last_query_ms = ms
        d2 = _bits_to_f32((keys >> np.uint64(32)).astype(np.uint32))
        idx = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32)
        return offs.astype(np.int64), d2.copy(), idx.copy()

    # -- updates --
    def _refit_items(self, items: np.ndarray, dirty, count_name: str) -> None:
        if items.shape[0] != self._items.shape[0]:
            raise ValueError(f"update must keep the {count_name} count; rebuild instead")
        ids = _ids(dirty)
        if ids is not None and ids.size and (ids.min() < 0 or ids.max() >= items.shape[0]):
            raise ValueError("dirty index out of range")
        self._refit = self._bvh.refit(items, ids)
        self._refit_ms = float(self._refit.total_ms)
        self._items = items

    def _update_topology(self, items: np.ndarray, deleted, moved, inserted) -> dict:
        d, m, ins = _ids(deleted), _ids(moved), _ids(inserted)
        nd = 0 if d is None else d.size
        ni = 0 if ins is None else ins.size
        appended = items.shape[0] - self._items.shape[0] + nd - ni
        if appended < 0 or appended > 2**31 - 1:
            raise ValueError("new primitive count must be old count - deleted + inserted + appended")
        t = self._bvh.update(items, d, m, ins, appended)  # invalid edits raise ValueError
        self._items = items
        self._wide = self._bvh.info().wide_count > 0
        self._refit_ms = float(t.refit_ms)
        return {"upload_ms": t.upload_ms, "plan_ms": t.plan_ms, "surgery_ms": t.surgery_ms,
                "relayout_ms": t.relayout_ms, "merge_ms": t.merge_ms, "wide_ms": t.wide_ms,
                "total_ms": t.total_ms, "refit_ms": t.refit_ms, "deleted": t.deleted,
                "moved": t.moved, "inserted": t.inserted, "appended": t.appended,
                "surgery_roots": t.surgery_roots, "rebuilt_leaves": t.rebuilt_leaves,
                "new_depth": t.new_depth, "full_rebuild": bool(t.full_rebuild)}


class Scene(_Index):
    """Static triangle mesh with a GPU BVH (quarry_py.cpp:907-968).

    Scene(triangles, builder='lbvh') with triangles of shape (M, 3, 3), or
    Scene(vertices, indices, builder='lbvh') with shared vertices (N, 3) and
    indices (M, 3)."""

    _kind = A.PRIM_TRI

    def __init__(self, *args, builder: str = "lbvh", **kw):
        self._indices = None
        if "triangles" in kw or (len(args) >= 1 and "vertices" not in kw and
                                 (len(args) == 1 or isinstance(args[1], str))):
            tris = kw.pop("triangles") if "triangles" in kw else args[0]
            if len(args) == 2:
                builder = args[1]
            items = _triangles(tris).reshape(-1, 9)
        else:
            vertices = kw.pop("vertices") if "vertices" in kw else args[0]
            indices = kw.pop("indices") if "indices" in kw else args[1]
            if len(args) == 3:
                builder = args[2]
            items = _mesh(vertices, indices).reshape(-1, 9)
            self._indices = np.ascontiguousarray(np.asarray(indices, dtype=np.int32))
        if kw:
            raise TypeError(f"unexpected arguments: {sorted(kw)}")
        self._build(np.ascontiguousarray(items), builder)

    @property
    def triangle_count(self) -> int:
        return int(self._items.shape[0])

    def update_triangles(self, triangles, dirty=None) -> None:
        self._refit_items(_triangles(triangles).reshape(-1, 9), dirty, "triangle")

    def update_vertices(self, vertices, dirty=None) -> None:
        if self._indices is None:
            raise ValueError("Scene was built from triangles; use update_triangles")
        self._refit_items(_mesh(vertices, self._indices).reshape(-1, 9), dirty, "triangle")

    def update_topology(self, triangles, deleted=None, moved=None, inserted=None) -> dict:
        out = self._update_topology(_triangles(triangles).reshape(-1, 9), deleted, moved, inserted)
        self._indices = None  # the mesh binding no longer describes the scene
        return out

    def raycast(self, origins, dirs, tmax: float = math.inf, order: str = "input",
                wide: bool = False):
        """Returns (hit bool[N], t float32[N]); t is inf on a miss (quarry_py.cpp:431-468)."""
        o = _points(origins, "origins")
        d = _points(dirs, "dirs")
        if o.shape[0] != d.shape[0]:
            raise ValueError("origins and dirs must match")
        rays = np.empty((o.shape[0], 8), np.float32)
        rays[:, 0:3] = o
        rays[:, 3:6] = d
        rays[:, 6] = np.float32(1e-6)
        rays[:, 7] = np.float32(tmax)
        res, ms, _ = self._bvh.raycast(rays, _order(order, _RAY_ORDERS), self._use_wide(wide))
        self._last_query_ms = ms
        hit = (res >> np.uint64(32)) != 0
        t = (res & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.float32).copy()
        t[~hit] = np.inf
        return hit, t


class PointCloud(_Index):
    """Static point cloud with a GPU BVH (quarry_py.cpp:970-1004)."""

    _kind = A.PRIM_POINT

    def __init__(self, points, builder: str = "lbvh"):
        p = _points(points)
        if p.shape[0] < 2:
            raise ValueError("at least 2 points are required")
        self._build(p, builder)

    @property
    def point_count(self) -> int:
        return int(self._items.shape[0])

    def update_points(self, points, dirty=None) -> None:
        self._refit_items(_points(points), dirty, "point")

    def update_topology(self, points, deleted=None, moved=None, inserted=None) -> dict:
        return self._update_topology(_points(points), deleted, moved, inserted)


class GaussianCloud(_Index):
    """Static Gaussian cloud with a GPU BVH; distances are to the solid ellipsoid
    (0 inside) (quarry_py.cpp:1006-1049)."""

    _kind = A.PRIM_GAUSS

    def __init__(self, means, rot, scale, builder: str = "lbvh"):
        g = _gaussians(means, rot, scale)
        if g.shape[0] < 2:
            raise ValueError("at least 2 gaussians are required")
        self._build(g, builder)

    @property
    def gaussian_count(self) -> int:
        return int(self._items.shape[0])

    def update(self, means, rot, scale, dirty=None) -> None:
        self._refit_items(_gaussians(means, rot, scale), dirty, "gaussian")

    def update_topology(self, means, rot, scale, deleted=None, moved=None, inserted=None) -> dict:
        return self._update_topology(_gaussians(means, rot, scale), deleted, moved, inserted)
