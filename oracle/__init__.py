"""CPU oracle for the BVH build + query path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package; the product never does.  It wraps liboracle.so (oracle.cpp, the
FP32 restatement of the reference's CPU path) and, when built,
_ref/libquarry_ref.so (the reference's own CPU sources compiled from
/root/reference with `make ref`).

Pinning (what checks each oracle function; see tests/test_oracle_*.py):
  * leaf math (rayTri, closestOnTri, boxDist2, gaussian distance/bounds): the
    reference self-test known answers (src/cpu/selftest.cpp:28-52, 100-123,
    355-398), an FP64 bisection witness (selftest.cpp:400-451), and bitwise
    equality with _ref on random inputs;
  * brute force / median-split BVH / kd-tree / ellipsoid BVH queries: brute
    force == tree bitwise, numpy float64 brute force within tolerance, and
    bitwise equality with _ref;
  * fingerprints: closed forms of src/core/fingerprint.h recomputed in numpy,
    and the reference's published hashes replayed through _ref generators;
  * CPU LBVH node bytes: structural invariants plus GPU tree-hash equality.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libquarry_ref.so")

PRIM_TRI, PRIM_POINT, PRIM_GAUSS = 0, 1, 2
ITEM_FLOATS = {PRIM_TRI: 9, PRIM_POINT: 3, PRIM_GAUSS: 10}

_lib = None
_ref = None

_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


def build(ref: bool = True) -> None:
    """Compiles liboracle.so (and _ref/ when /root/reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        L = ctypes.CDLL(LIB)
        L.orc_ray_brute.argtypes = [_f32p, ctypes.c_int, _f32p, ctypes.c_int, _u64p, ctypes.c_int]
        L.orc_knn_brute.argtypes = [ctypes.c_int, _vp, ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int, _u32p, ctypes.c_int]
        L.orc_radius_brute.argtypes = [ctypes.c_int, _vp, ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_float, _i32p, _i32p, _u64p, ctypes.c_int]
        L.orc_bvh_new.argtypes = [ctypes.c_int, _vp, ctypes.c_int]
        L.orc_bvh_new.restype = _vp
        L.orc_bvh_free.argtypes = [_vp]
        L.orc_bvh_ray.argtypes = [_vp, _f32p, ctypes.c_int, _u64p, ctypes.c_int]
        L.orc_bvh_knn.argtypes = [_vp, _f32p, ctypes.c_int, ctypes.c_int, _u32p, ctypes.c_int]
        L.orc_bvh_radius.argtypes = [_vp, _f32p, ctypes.c_int, ctypes.c_float, _i32p, _i32p, _u64p, ctypes.c_int]
        L.orc_kd_new.argtypes = [_f32p, ctypes.c_int]
        L.orc_kd_new.restype = _vp
        L.orc_kd_free.argtypes = [_vp]
        L.orc_kd_knn.argtypes = [_vp, _f32p, ctypes.c_int, ctypes.c_int, _u32p, ctypes.c_int]
        L.orc_kd_radius.argtypes = [_vp, _f32p, ctypes.c_int, ctypes.c_float, _i32p, _i32p, _u64p, ctypes.c_int]
        L.orc_lbvh_nodes.argtypes = [ctypes.c_int, _vp, ctypes.c_int, _vp, _i32p]
        L.orc_lbvh_nodes.restype = ctypes.c_int
        L.orc_fnv_bytes.argtypes = [_vp, ctypes.c_size_t]
        L.orc_fnv_bytes.restype = ctypes.c_uint64
        L.orc_ray_hash.argtypes = [_u64p, ctypes.c_int, _u64p, ctypes.POINTER(ctypes.c_double)]
        L.orc_ray_hash.restype = ctypes.c_uint64
        L.orc_closest_hash.argtypes = [_u32p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.orc_closest_hash.restype = ctypes.c_uint64
        L.orc_knn_hash.argtypes = [_u32p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.orc_knn_hash.restype = ctypes.c_uint64
        L.orc_csr_hash.argtypes = [_i32p, _u64p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.orc_csr_hash.restype = ctypes.c_uint64
        L.orc_tri_intersect.argtypes = [_f32p, _f32p, _f32p]
        L.orc_tri_closest.argtypes = [_f32p, _f32p, _i32p, _f32p, _f32p]
        L.orc_tri_closest.restype = ctypes.c_float
        L.orc_gauss_dist2.argtypes = [_f32p, _f32p, _f32p, _i32p]
        L.orc_gauss_dist2.restype = ctypes.c_float
        L.orc_gauss_bounds.argtypes = [_f32p, _f32p]
        L.orc_box_dist2.argtypes = [_f32p, _f32p]
        L.orc_box_dist2.restype = ctypes.c_float
        L.orc_ray_box.argtypes = [_f32p, _f32p, ctypes.c_float, _f32p]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref():
    """The reference's own CPU implementation (None if _ref was not built)."""
    global _ref
    if _ref is None:
        if not ref_available():
            return None
        R = ctypes.CDLL(REF_LIB)
        R.ref_selftest.restype = ctypes.c_int
        R.ref_scene.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, _f32p, ctypes.c_int, _f32p]
        R.ref_scene.restype = ctypes.c_int
        R.ref_rays.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, _f32p]
        R.ref_query_points.argtypes = [_f32p, ctypes.c_int, ctypes.c_ulonglong, _f32p]
        for name in ("bvh", "kd", "gauss"):
            getattr(R, f"ref_{name}_new").argtypes = [_f32p, ctypes.c_int]
            getattr(R, f"ref_{name}_new").restype = _vp
            getattr(R, f"ref_{name}_free").argtypes = [_vp]
            getattr(R, f"ref_{name}_knn").argtypes = [_vp, _f32p, ctypes.c_int, ctypes.c_int, _u32p, ctypes.c_int]
            getattr(R, f"ref_{name}_radius").argtypes = [_vp, _f32p, ctypes.c_int, ctypes.c_float, _i32p, _i32p, _u64p, ctypes.c_int]
        R.ref_bvh_ray.argtypes = [_vp, _f32p, ctypes.c_int, _u64p, ctypes.c_int]
        R.ref_gauss_dist2.argtypes = [_f32p, _f32p]
        R.ref_gauss_dist2.restype = ctypes.c_float
        R.ref_tri_closest.argtypes = [_f32p, _f32p]
        R.ref_tri_closest.restype = ctypes.c_float
        R.ref_tri_intersect.argtypes = [_f32p, _f32p, _f32p]
        R.ref_tri_intersect.restype = ctypes.c_int
        _ref = R
    return _ref


LEARN_LIB = os.path.join(HERE, "_ref", "liblbvh_learn.so")
_learn = None


def learn_lbvh(tris):
    """The reference's LBVH walkthrough (learn/03_lbvh.cpp, compiled by `make ref`)
    on triangles: (leaf order, left (n-1), right (n-1), depth), or None when that
    build is absent."""
    global _learn
    if _learn is None:
        if not os.path.exists(LEARN_LIB):
            return None
        L = ctypes.CDLL(LEARN_LIB)
        L.learn_lbvh.argtypes = [_f32p, ctypes.c_int, _i32p, _i32p, _i32p]
        L.learn_lbvh.restype = ctypes.c_int
        _learn = L
    t = _c32(tris).reshape(-1, 9)
    n = t.shape[0]
    order, left, right = (np.zeros(max(m, 1), np.int32) for m in (n, n - 1, n - 1))
    depth = _learn.learn_lbvh(_ptr(t, _f32p), n, _ptr(order, _i32p), _ptr(left, _i32p), _ptr(right, _i32p))
    return order[:n], left[:n - 1], right[:n - 1], depth


def _c32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _records(n: int, k: int) -> np.ndarray:
    return np.zeros((n, max(k, 1), 2), dtype=np.uint32)


# ---- brute force ---------------------------------------------------------------------

def ray_brute(tris, rays, threads: int = 1) -> np.ndarray:
    t, r = _c32(tris), _c32(rays)
    out = np.zeros(r.shape[0], dtype=np.uint64)
    lib().orc_ray_brute(_ptr(t, _f32p), t.shape[0], _ptr(r, _f32p), r.shape[0], _ptr(out, _u64p), threads)
    return out


def knn_brute(kind: int, items, pts, k: int, threads: int = 1) -> np.ndarray:
    """k = 0: closest.  Returns uint32 records (n, max(k,1), 2): dist2 bits, index."""
    it, p = _c32(items), _c32(pts)
    out = _records(p.shape[0], k)
    lib().orc_knn_brute(kind, it.ctypes.data, it.shape[0], _ptr(p, _f32p), p.shape[0], k, _ptr(out, _u32p), threads)
    return out


def _csr(fn, *args, nq: int):
    counts = np.zeros(nq, dtype=np.int32)
    fn(*args, _ptr(counts, _i32p), None, None)
    offsets = np.zeros(nq + 1, dtype=np.int32)
    np.cumsum(counts, out=offsets[1:])
    keys = np.zeros(max(int(offsets[-1]), 1), dtype=np.uint64)
    fn(*args, _ptr(counts, _i32p), _ptr(offsets, _i32p), _ptr(keys, _u64p))
    return offsets, keys[: int(offsets[-1])]


def radius_brute(kind: int, items, pts, r: float, threads: int = 1):
    it, p = _c32(items), _c32(pts)
    L = lib()

    def fn(counts, offs, keys):
        L.orc_radius_brute(kind, it.ctypes.data, it.shape[0], _ptr(p, _f32p), p.shape[0], r, counts, offs, keys, threads)

    return _csr(fn, nq=p.shape[0])


# ---- trees -----------------------------------------------------------------------------

class CpuBvh:
    """Median-split BVH, leaf 4 (src/cpu/bvh.cpp, src/cpu/gauss_bvh.cpp)."""

    def __init__(self, kind: int, items):
        self.kind = kind
        self._items = _c32(items)
        self.h = lib().orc_bvh_new(kind, self._items.ctypes.data, self._items.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bvh_free(self.h)
            self.h = None

    def ray(self, rays, threads: int = 1) -> np.ndarray:
        r = _c32(rays)
        out = np.zeros(r.shape[0], dtype=np.uint64)
        lib().orc_bvh_ray(self.h, _ptr(r, _f32p), r.shape[0], _ptr(out, _u64p), threads)
        return out

    def knn(self, pts, k: int, threads: int = 1) -> np.ndarray:
        p = _c32(pts)
        out = _records(p.shape[0], k)
        lib().orc_bvh_knn(self.h, _ptr(p, _f32p), p.shape[0], k, _ptr(out, _u32p), threads)
        return out

    def radius(self, pts, r: float, threads: int = 1):
        p = _c32(pts)

        def fn(counts, offs, keys):
            lib().orc_bvh_radius(self.h, _ptr(p, _f32p), p.shape[0], r, counts, offs, keys, threads)

        return _csr(fn, nq=p.shape[0])


class CpuKd:
    """Point kd-tree, leaf 8 (src/cpu/point_kd.cpp)."""

    def __init__(self, pts):
        self._pts = _c32(pts)
        self.h = lib().orc_kd_new(_ptr(self._pts, _f32p), self._pts.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_kd_free(self.h)
            self.h = None

    def knn(self, pts, k: int, threads: int = 1) -> np.ndarray:
        p = _c32(pts)
        out = _records(p.shape[0], k)
        lib().orc_kd_knn(self.h, _ptr(p, _f32p), p.shape[0], k, _ptr(out, _u32p), threads)
        return out

    def radius(self, pts, r: float, threads: int = 1):
        p = _c32(pts)

        def fn(counts, offs, keys):
            lib().orc_kd_radius(self.h, _ptr(p, _f32p), p.shape[0], r, counts, offs, keys, threads)

        return _csr(fn, nq=p.shape[0])


def lbvh_nodes(kind: int, items):
    """CPU LBVH in the GPU node layout: (nodes bytes (2n-1, 32) as uint8, prim_ids, depth)."""
    it = _c32(items)
    n = it.shape[0]
    nodes = np.zeros((2 * n - 1) * 32, dtype=np.uint8)
    ids = np.zeros(n, dtype=np.int32)
    depth = lib().orc_lbvh_nodes(kind, it.ctypes.data, n, nodes.ctypes.data, _ptr(ids, _i32p))
    return nodes, ids, depth


def fnv_bytes(buf: np.ndarray) -> int:
    b = np.ascontiguousarray(buf)
    return int(lib().orc_fnv_bytes(b.ctypes.data, b.nbytes))


# ---- fingerprints (src/core/fingerprint.h) ------------------------------------------

def ray_hash(results: np.ndarray):
    v = np.ascontiguousarray(results, dtype=np.uint64)
    hits = ctypes.c_uint64()
    s = ctypes.c_double()
    h = lib().orc_ray_hash(_ptr(v, _u64p), v.shape[0], ctypes.byref(hits), ctypes.byref(s))
    return int(h), int(hits.value), float(s.value)


def closest_hash(records: np.ndarray):
    r = np.ascontiguousarray(records, dtype=np.uint32).reshape(-1, 2)
    s = ctypes.c_double()
    h = lib().orc_closest_hash(_ptr(r, _u32p), r.shape[0], ctypes.byref(s))
    return int(h), float(s.value)


def knn_hash(records: np.ndarray, k: int):
    r = np.ascontiguousarray(records, dtype=np.uint32).reshape(-1, k, 2)
    s = ctypes.c_double()
    h = lib().orc_knn_hash(_ptr(r, _u32p), r.shape[0], k, ctypes.byref(s))
    return int(h), float(s.value)


def csr_hash(offsets: np.ndarray, keys: np.ndarray):
    o = np.ascontiguousarray(offsets, dtype=np.int32)
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    if k.size == 0:
        k = np.zeros(1, dtype=np.uint64)
    s = ctypes.c_double()
    h = lib().orc_csr_hash(_ptr(o, _i32p), _ptr(k, _u64p), o.shape[0] - 1, ctypes.byref(s))
    return int(h), float(s.value)


# ---- leaf math probes --------------------------------------------------------------------

def tri_intersect(ray, tri):
    r, t = _c32(ray), _c32(tri)
    out = np.zeros(1, dtype=np.float32)
    ok = lib().orc_tri_intersect(_ptr(r, _f32p), _ptr(t, _f32p), _ptr(out, _f32p))
    return bool(ok), float(out[0])


def tri_closest(p, tri):
    pp, t = _c32(p), _c32(tri)
    reg = np.zeros(1, dtype=np.int32)
    b = np.zeros(2, dtype=np.float32)
    d = lib().orc_tri_closest(_ptr(pp, _f32p), _ptr(t, _f32p), _ptr(reg, _i32p), _ptr(b[:1], _f32p), _ptr(b[1:], _f32p))
    return np.float32(d), int(reg[0]), b


def gauss_dist2(p, g):
    pp, gg = _c32(p), _c32(g)
    t = np.zeros(1, dtype=np.float32)
    inside = np.zeros(1, dtype=np.int32)
    d = lib().orc_gauss_dist2(_ptr(pp, _f32p), _ptr(gg, _f32p), _ptr(t, _f32p), _ptr(inside, _i32p))
    return np.float32(d), float(t[0]), bool(inside[0])


def gauss_bounds(g) -> np.ndarray:
    gg = _c32(g)
    out = np.zeros(6, dtype=np.float32)
    lib().orc_gauss_bounds(_ptr(gg, _f32p), _ptr(out, _f32p))
    return out


def box_dist2(p, box6) -> np.float32:
    pp, b = _c32(p), _c32(box6)
    return np.float32(lib().orc_box_dist2(_ptr(pp, _f32p), _ptr(b, _f32p)))
