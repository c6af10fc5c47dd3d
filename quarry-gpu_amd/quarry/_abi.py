"""ctypes binding of libquarry_hip.so (include/quarry/quarry_abi.h).

This is the only place Python touches the native library.  Every call checks
the returned status and raises.  The reference maps std::runtime_error (bad
query parameters, device failures) to RuntimeError and std::invalid_argument
(bad topology edits) to ValueError (docs/api.md:390-392), so the mapping here is

    QR_ERR_INVALID, QR_ERR_DEVICE, QR_ERR_OOM, QR_ERR_NO_DEVICE -> QuarryError
        (a RuntimeError subclass)
    QR_ERR_EDIT -> ValueError

Argument validation that the reference's Python layer performs up front
(shapes, k, radius) raises ValueError in quarry/__init__.py before any call.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libquarry_hip.so")

QR_OK, QR_ERR_INVALID, QR_ERR_DEVICE, QR_ERR_OOM, QR_ERR_EDIT, QR_ERR_NO_DEVICE = range(6)
PRIM_TRI, PRIM_POINT, PRIM_GAUSS = 0, 1, 2
BUILDER_LBVH, BUILDER_SAH = 0, 1
ORDER_INPUT, ORDER_MORTON, ORDER_MORTON_DIRECTION = 0, 1, 2
ITEM_FLOATS = {PRIM_TRI: 9, PRIM_POINT: 3, PRIM_GAUSS: 10}


class QuarryError(RuntimeError):
    pass


class BuildTimings(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "morton_ms", "sort_ms", "tree_ms", "scatter_ms", "refit_ms",
        "sah_bin_ms", "sah_split_ms", "sah_part_ms", "total_ms")]


class RefitTimings(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "upload_ms", "gather_ms", "tree_ms", "mark_ms", "merge_ms", "wide_ms", "total_ms")] + [
        ("updated", ctypes.c_int)]


class TopoTimings(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("upload_ms", "plan_ms", "surgery_ms", "relayout_ms",
                                                "merge_ms", "wide_ms", "total_ms")] + \
               [(n, ctypes.c_int) for n in ("deleted", "moved", "inserted", "appended", "surgery_roots",
                                             "rebuilt_leaves", "new_depth", "full_rebuild")] + \
               [("refit_ms", ctypes.c_double)]


class TraversalStats(ctypes.Structure):
    _fields_ = [("node_tests", ctypes.c_uint64), ("leaf_tests", ctypes.c_uint64),
                ("stack_overflows", ctypes.c_uint64), ("sort_ms", ctypes.c_double),
                ("kernel_ms", ctypes.c_double), ("nodes_opened", ctypes.c_uint64)]


class RayStats(ctypes.Structure):
    _fields_ = [("rays", ctypes.c_uint64), ("hits", ctypes.c_uint64), ("hash", ctypes.c_uint64),
                ("sum_t", ctypes.c_double)]


class SahStats(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int), ("binned_slots", ctypes.c_longlong),
                ("partitioned_slots", ctypes.c_longlong), ("serial_slots", ctypes.c_longlong)]


class BvhInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "prim_count", "node_count", "depth", "wide_count", "prim_kind", "builder")] + [
        ("d_nodes", ctypes.c_void_p), ("d_items", ctypes.c_void_p),
        ("d_prim_ids", ctypes.c_void_p), ("d_wide_nodes", ctypes.c_void_p)]


_lib = None
_vp = ctypes.c_void_p
_ip = ctypes.c_int
_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QuarryError(
                f"{LIB_PATH} is missing: build it with `python quarry-gpu_amd/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "qr_last_error": ([], ctypes.c_char_p),
            "qr_describe": ([ctypes.c_char_p, _ip], _ip),
            "qr_warmup": ([], _ip),
            "qr_set_stream": ([_vp], _ip),
            "qr_dev_alloc": ([ctypes.c_size_t, _vp], _ip),
            "qr_dev_free": ([_vp], _ip),
            "qr_dev_upload": ([_vp, _vp, ctypes.c_size_t], _ip),
            "qr_dev_download": ([_vp, _vp, ctypes.c_size_t], _ip),
            "qr_dev_synchronize": ([], _ip),
            "qr_dev_copy": ([_vp, _vp, ctypes.c_size_t, _dp], _ip),
            "qr_set_traversal_mode": ([_ip], _ip),
            "qr_set_frontier_cap": ([_ip], _ip),
            "qr_set_build_phase_timing": ([_ip], _ip),
            "qr_bvh_build": ([_ip, _ip, _vp, _ip, _vp, _vp], _ip),
            "qr_bvh_build_device": ([_ip, _ip, _vp, _ip, _vp, _vp], _ip),
            "qr_bvh_destroy": ([_vp], _ip),
            "qr_bvh_get_info": ([_vp, _vp], _ip),
            "qr_bvh_build_wide": ([_vp, _dp], _ip),
            "qr_bvh_refit": ([_vp, _vp, _ip, _vp, _ip, _vp, _dp], _ip),
            "qr_bvh_refit_device": ([_vp, _vp, _ip, _vp, _dp], _ip),
            "qr_bvh_validate": ([_vp, _vp, _ip, ctypes.c_char_p, _ip], _ip),
            "qr_bvh_tree_hash": ([_vp, ctypes.POINTER(ctypes.c_uint64)], _ip),
            "qr_bvh_read": ([_vp, _vp, _vp], _ip),
            "qr_rays_upload": ([_vp, _ip, _vp], _ip),
            "qr_rays_destroy": ([_vp], _ip),
            "qr_raycast": ([_vp, _vp, _ip, _ip, _ip, _dp], _ip),
            "qr_rays_summarize": ([_vp, _vp, _vp], _ip),
            "qr_last_traversal_stats": ([_vp], _ip),
            "qr_raycast_device": ([_vp, _vp, _ip, _vp, _ip, _ip, _dp], _ip),
            "qr_points_upload": ([_vp, _ip, _ip, _vp], _ip),
            "qr_points_destroy": ([_vp], _ip),
            "qr_closest": ([_vp, _vp, _ip, _ip, _ip, _dp], _ip),
            "qr_knn": ([_vp, _vp, _ip, _ip, _ip, _ip, _dp], _ip),
            "qr_radius": ([_vp, _vp, ctypes.c_float, _ip, _ip, _ip, _vp, _dp], _ip),
            "qr_points_read": ([_vp, _vp], _ip),
            "qr_points_info": ([_vp, _vp, _vp], _ip),
            "qr_radius_info": ([_vp, _vp, _vp], _ip),
            "qr_radius_read": ([_vp, _vp, _vp], _ip),
            "qr_radius_destroy": ([_vp], _ip),
            "qr_last_query_stats": ([_vp], _ip),
            "qr_closest_device": ([_vp, _vp, _ip, _vp, _ip, _ip, _dp], _ip),
            "qr_knn_device": ([_vp, _vp, _ip, _ip, _vp, _ip, _ip, _dp], _ip),
            "qr_knn_device_noalloc": ([_vp, _vp, _ip, _ip, _vp, _ip, _ip], _ip),
            "qr_last_sah_stats": ([ctypes.POINTER(SahStats)], _ip),
            "qr_bvh_update": ([_vp, _vp, _ip, _vp, _ip, _vp, _ip, _vp, _ip, _ip, _vp, _dp], _ip),
            "qr_rays_device_ptrs": ([_vp, _vp, _vp, _vp], _ip),
            "qr_points_device_ptrs": ([_vp, _vp, _vp], _ip),
            "qr_radius_device_ptrs": ([_vp, _vp, _vp], _ip),
            "qr_batch_build": ([_ip, _vp, _vp, _ip, _vp, _vp], _ip),
            "qr_batch_build_device": ([_ip, _vp, _vp, _ip, _vp, _vp], _ip),
            "qr_batch_destroy": ([_vp], _ip),
            "qr_batch_info": ([_vp, _vp, _vp, _vp], _ip),
            "qr_batch_read_scene": ([_vp, _ip, _vp, _vp], _ip),
            "qr_batch_tree_hash": ([_vp, _ip, _vp], _ip),
            "qr_batch_plan_create": ([_vp, _vp, _vp, _ip, _ip, _vp], _ip),
            "qr_batch_plan_run": ([_vp, _dp], _ip),
            "qr_batch_plan_run_eager": ([_vp, _dp], _ip),
            "qr_batch_plan_read": ([_vp, _vp], _ip),
            "qr_batch_plan_config": ([_vp, _vp, _vp, _dp], _ip),
            "qr_batch_plan_destroy": ([_vp], _ip),
            "qr_raycast_save_device": ([_vp, _vp, _ip, _vp, _vp, _ip, _ip, _dp], _ip),
            "qr_closest_save_aux": ([_vp, _vp, _ip, _vp, _vp], _ip),
            "qr_adjoint_closest": ([_vp, _vp, _ip, _vp, _vp, _vp, _vp, _vp], _ip),
            "qr_adjoint_ray": ([_vp, _vp, _ip, _vp, _vp, _vp, _vp, _vp, _vp], _ip),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return [n for n in dir(lib()) if n.startswith("qr_")]


def check(status: int) -> None:
    if status == QR_OK:
        return
    msg = lib().qr_last_error().decode(errors="replace")
    if status == QR_ERR_EDIT:
        raise ValueError(msg)
    raise QuarryError(msg)


def ptr(a: np.ndarray | None):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def describe() -> str | None:
    buf = ctypes.create_string_buffer(256)
    st = lib().qr_describe(buf, 256)
    return buf.value.decode() if st == QR_OK else None


class Bvh:
    """Owning wrapper of a qr_bvh handle."""

    def __init__(self, kind: int, items: np.ndarray, builder: int = BUILDER_LBVH):
        items = np.ascontiguousarray(items, dtype=np.float32).reshape(-1, ITEM_FLOATS[kind])
        self.kind = kind
        self.h = ctypes.c_void_p()
        self.timings = BuildTimings()
        check(lib().qr_bvh_build(kind, builder, ptr(items), items.shape[0],
                                 ctypes.byref(self.timings), ctypes.byref(self.h)))

    @classmethod
    def from_device(cls, kind: int, dev_ptr: int, n: int, builder: int = BUILDER_LBVH) -> "Bvh":
        self = cls.__new__(cls)
        self.kind = kind
        self.h = ctypes.c_void_p()
        self.timings = BuildTimings()
        check(lib().qr_bvh_build_device(kind, builder, ctypes.c_void_p(dev_ptr), n,
                                        ctypes.byref(self.timings), ctypes.byref(self.h)))
        return self

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            check(lib().qr_bvh_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> BvhInfo:
        out = BvhInfo()
        check(lib().qr_bvh_get_info(self.h, ctypes.byref(out)))
        return out

    def build_wide(self) -> float:
        ms = ctypes.c_double()
        check(lib().qr_bvh_build_wide(self.h, ctypes.byref(ms)))
        return ms.value

    def refit(self, items: np.ndarray, dirty: np.ndarray | None = None) -> RefitTimings:
        items = np.ascontiguousarray(items, dtype=np.float32).reshape(-1, ITEM_FLOATS[self.kind])
        t = RefitTimings()
        ms = ctypes.c_double()
        if dirty is None:
            check(lib().qr_bvh_refit(self.h, ptr(items), items.shape[0], None, 0, ctypes.byref(t), ctypes.byref(ms)))
        else:
            d = np.ascontiguousarray(dirty, dtype=np.int32)
            check(lib().qr_bvh_refit(self.h, ptr(items), items.shape[0], ptr(d), d.shape[0], ctypes.byref(t), ctypes.byref(ms)))
        return t

    def refit_device(self, dev_ptr: int, n: int) -> float:
        """Full refit from device-resident scene-order items; returns device ms."""
        t = RefitTimings()
        ms = ctypes.c_double()
        check(lib().qr_bvh_refit_device(self.h, ctypes.c_void_p(dev_ptr), n, ctypes.byref(t), ctypes.byref(ms)))
        return ms.value

    def update(self, items: np.ndarray, deleted=None, moved=None, inserted=None,
               appended: int = 0) -> TopoTimings:
        """Topology edit: items is the whole new scene (see qr_bvh_update)."""
        items = np.ascontiguousarray(items, dtype=np.float32).reshape(-1, ITEM_FLOATS[self.kind])
        arrs = [np.ascontiguousarray(a if a is not None else [], dtype=np.int32)
                for a in (deleted, moved, inserted)]
        t = TopoTimings()
        ms = ctypes.c_double()
        check(lib().qr_bvh_update(self.h, ptr(items), items.shape[0], ptr(arrs[0]), arrs[0].shape[0],
                                  ptr(arrs[1]), arrs[1].shape[0], ptr(arrs[2]), arrs[2].shape[0],
                                  int(appended), ctypes.byref(t), ctypes.byref(ms)))
        return t

    def validate(self, items: np.ndarray) -> str:
        items = np.ascontiguousarray(items, dtype=np.float32).reshape(-1, ITEM_FLOATS[self.kind])
        buf = ctypes.create_string_buffer(512)
        check(lib().qr_bvh_validate(self.h, ptr(items), items.shape[0], buf, 512))
        return buf.value.decode()

    def tree_hash(self) -> int:
        out = ctypes.c_uint64()
        check(lib().qr_bvh_tree_hash(self.h, ctypes.byref(out)))
        return int(out.value)

    def read(self):
        inf = self.info()
        nodes = np.zeros(inf.node_count * 32, dtype=np.uint8)
        ids = np.zeros(inf.prim_count, dtype=np.int32)
        check(lib().qr_bvh_read(self.h, ptr(nodes), ptr(ids)))
        return nodes, ids

    # ---- queries -------------------------------------------------------------------
    def raycast(self, rays: np.ndarray, order: int = ORDER_INPUT, wide: bool = False,
                stats: bool = False):
        """Returns (u64 results, ms, stats or None)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        rb = ctypes.c_void_p()
        check(lib().qr_rays_upload(ptr(rays), rays.shape[0], ctypes.byref(rb)))
        try:
            ms = ctypes.c_double()
            check(lib().qr_raycast(self.h, rb, int(stats), order, int(wide), ctypes.byref(ms)))
            out = np.zeros(rays.shape[0], dtype=np.uint64)
            rs = RayStats()
            check(lib().qr_rays_summarize(rb, ptr(out), ctypes.byref(rs)))
            st = TraversalStats()
            check(lib().qr_last_traversal_stats(ctypes.byref(st)))
            return out, ms.value, st
        finally:
            lib().qr_rays_destroy(rb)

    def _points(self, pts: np.ndarray, stride: int):
        pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
        pb = ctypes.c_void_p()
        check(lib().qr_points_upload(ptr(pts), pts.shape[0], stride, ctypes.byref(pb)))
        return pts, pb

    def closest(self, pts: np.ndarray, order: int = ORDER_INPUT, wide: bool = False,
                stats: bool = False):
        """Returns (records uint32 (n, 1, 2), ms, stats or None)."""
        pts, pb = self._points(pts, 1)
        try:
            ms = ctypes.c_double()
            check(lib().qr_closest(self.h, pb, int(stats), order, int(wide), ctypes.byref(ms)))
            out = np.zeros((pts.shape[0], 1, 2), dtype=np.uint32)
            check(lib().qr_points_read(pb, ptr(out)))
            return out, ms.value, self._qstats(stats)
        finally:
            lib().qr_points_destroy(pb)

    def knn(self, pts: np.ndarray, k: int, order: int = ORDER_INPUT, wide: bool = False,
            stats: bool = False):
        pts, pb = self._points(pts, k)
        try:
            ms = ctypes.c_double()
            check(lib().qr_knn(self.h, pb, k, int(stats), order, int(wide), ctypes.byref(ms)))
            out = np.zeros((pts.shape[0], k, 2), dtype=np.uint32)
            check(lib().qr_points_read(pb, ptr(out)))
            return out, ms.value, self._qstats(stats)
        finally:
            lib().qr_points_destroy(pb)

    def radius(self, pts: np.ndarray, r: float, order: int = ORDER_INPUT, wide: bool = False,
               stats: bool = False):
        """Returns (offsets int32 (n+1), keys uint64, ms, stats or None)."""
        pts, pb = self._points(pts, 1)
        csr = ctypes.c_void_p()
        try:
            ms = ctypes.c_double()
            check(lib().qr_radius(self.h, pb, float(r), int(stats), order, int(wide),
                                  ctypes.byref(csr), ctypes.byref(ms)))
            cnt, tot = ctypes.c_int(), ctypes.c_int()
            check(lib().qr_radius_info(csr, ctypes.byref(cnt), ctypes.byref(tot)))
            offs = np.zeros(cnt.value + 1, dtype=np.int32)
            keys = np.zeros(max(tot.value, 1), dtype=np.uint64)
            check(lib().qr_radius_read(csr, ptr(offs), ptr(keys)))
            return offs, keys[: tot.value], ms.value, self._qstats(stats)
        finally:
            if csr.value:
                lib().qr_radius_destroy(csr)
            lib().qr_points_destroy(pb)

    @staticmethod
    def _qstats(want: bool):
        st = TraversalStats()
        check(lib().qr_last_query_stats(ctypes.byref(st)))
        return st


class PointBuffer:
    """Device-resident query points + result slots (uploadPointsGPU)."""

    def __init__(self, pts: np.ndarray, stride: int = 1):
        pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
        self.count = pts.shape[0]
        self.stride = stride
        self.h = ctypes.c_void_p()
        check(lib().qr_points_upload(ptr(pts), self.count, stride, ctypes.byref(self.h)))

    def close(self) -> None:
        if self.h.value:
            lib().qr_points_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def closest(self, bvh: Bvh, order: int = ORDER_MORTON, wide: bool = True, stats: bool = False) -> float:
        ms = ctypes.c_double()
        check(lib().qr_closest(bvh.h, self.h, int(stats), order, int(wide), ctypes.byref(ms)))
        return ms.value

    def knn(self, bvh: Bvh, k: int, order: int = ORDER_MORTON, wide: bool = True, stats: bool = False) -> float:
        ms = ctypes.c_double()
        check(lib().qr_knn(bvh.h, self.h, k, int(stats), order, int(wide), ctypes.byref(ms)))
        return ms.value

    def radius(self, bvh: Bvh, r: float, order: int = ORDER_MORTON, wide: bool = True,
               stats: bool = False, keep: bool = False):
        """Runs a radius query; returns ms (and the CSR arrays when keep=True)."""
        csr = ctypes.c_void_p()
        ms = ctypes.c_double()
        check(lib().qr_radius(bvh.h, self.h, float(r), int(stats), order, int(wide),
                              ctypes.byref(csr), ctypes.byref(ms)))
        try:
            if not keep:
                return ms.value
            cnt, tot = ctypes.c_int(), ctypes.c_int()
            check(lib().qr_radius_info(csr, ctypes.byref(cnt), ctypes.byref(tot)))
            offs = np.zeros(cnt.value + 1, dtype=np.int32)
            keys = np.zeros(max(tot.value, 1), dtype=np.uint64)
            check(lib().qr_radius_read(csr, ptr(offs), ptr(keys)))
            return ms.value, offs, keys[: tot.value]
        finally:
            lib().qr_radius_destroy(csr)

    def read(self) -> np.ndarray:
        out = np.zeros((self.count, self.stride, 2), dtype=np.uint32)
        check(lib().qr_points_read(self.h, ptr(out)))
        return out


class RayBuffer:
    """Device-resident rays + per-ray results (uploadRaysGPU)."""

    def __init__(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        self.count = rays.shape[0]
        self.h = ctypes.c_void_p()
        check(lib().qr_rays_upload(ptr(rays), self.count, ctypes.byref(self.h)))

    def close(self) -> None:
        if self.h.value:
            lib().qr_rays_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cast(self, bvh: Bvh, order: int = ORDER_MORTON, wide: bool = True, stats: bool = False) -> float:
        ms = ctypes.c_double()
        check(lib().qr_raycast(bvh.h, self.h, int(stats), order, int(wide), ctypes.byref(ms)))
        return ms.value

    def summarize(self):
        out = np.zeros(self.count, dtype=np.uint64)
        rs = RayStats()
        check(lib().qr_rays_summarize(self.h, ptr(out), ctypes.byref(rs)))
        return out, rs


def last_query_stats() -> TraversalStats:
    st = TraversalStats()
    check(lib().qr_last_query_stats(ctypes.byref(st)))
    return st


def last_sah_stats() -> SahStats:
    st = SahStats()
    check(lib().qr_last_sah_stats(ctypes.byref(st)))
    return st


def last_ray_stats() -> TraversalStats:
    st = TraversalStats()
    check(lib().qr_last_traversal_stats(ctypes.byref(st)))
    return st


class DeviceArray:
    """A device allocation holding a copy of a host array (for *_device calls)."""

    @classmethod
    def empty(cls, shape, dtype) -> "DeviceArray":
        self = cls.__new__(cls)
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        check(lib().qr_dev_alloc(max(self.nbytes, 1), ctypes.byref(self.ptr)))
        return self

    def at(self, byte_offset: int) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.ptr.value + int(byte_offset))

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.nbytes = host.nbytes
        self.shape = host.shape
        self.dtype = host.dtype
        self.ptr = ctypes.c_void_p()
        check(lib().qr_dev_alloc(self.nbytes, ctypes.byref(self.ptr)))
        check(lib().qr_dev_upload(self.ptr, ptr(host), self.nbytes))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().qr_dev_download(ptr(out), self.ptr, self.nbytes))
        return out

    def close(self) -> None:
        if self.ptr.value:
            lib().qr_dev_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Batched traversal variants (qr_batch_plan_config): binary or 4-wide nodes x the
# traversal stack in LDS (256- / 64-lane workgroups), scratch, or registers.
BATCH_VARIANTS = ("lds256", "lds64", "scratch64", "reg64", "wide_lds256", "wide_lds64", "wide_scratch64",
                  "wide_reg64")


class Batch:
    """Owning wrapper of a qr_batch: many scenes of one kind built together."""

    def __init__(self, kind: int, items: np.ndarray, offsets: np.ndarray):
        items = np.ascontiguousarray(items, dtype=np.float32).reshape(-1, ITEM_FLOATS[kind])
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        self.kind = kind
        self.h = ctypes.c_void_p()
        self.timings = BuildTimings()
        check(lib().qr_batch_build(kind, ptr(items), ptr(self.offsets), self.offsets.shape[0] - 1,
                                   ctypes.byref(self.timings), ctypes.byref(self.h)))

    @classmethod
    def from_device(cls, kind: int, dev_ptr: int, offsets: np.ndarray) -> "Batch":
        self = cls.__new__(cls)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        self.kind = kind
        self.h = ctypes.c_void_p()
        self.timings = BuildTimings()
        check(lib().qr_batch_build_device(kind, ctypes.c_void_p(dev_ptr), ptr(self.offsets),
                                          self.offsets.shape[0] - 1, ctypes.byref(self.timings),
                                          ctypes.byref(self.h)))
        return self

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            lib().qr_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        s, t, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().qr_batch_info(self.h, ctypes.byref(s), ctypes.byref(t), ctypes.byref(d)))
        return s.value, t.value, d.value

    def tree_hash(self, scene: int) -> int:
        out = ctypes.c_uint64()
        check(lib().qr_batch_tree_hash(self.h, scene, ctypes.byref(out)))
        return int(out.value)

    def read_scene(self, scene: int):
        n = int(self.offsets[scene + 1] - self.offsets[scene])
        nodes = np.zeros((2 * n - 1) * 32, np.uint8)
        ids = np.zeros(n, np.int32)
        check(lib().qr_batch_read_scene(self.h, scene, ptr(nodes), ptr(ids)))
        return nodes, ids


class BatchPlan:
    """A prepared batched query (device copy of the points, captured graph)."""

    def __init__(self, batch: Batch, points: np.ndarray, qoffsets: np.ndarray, k: int, tune=True):
        """tune: True autotunes, False takes the default, or (variant, morton) forces one."""
        self.batch = batch
        self.points = DeviceArray(np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3))
        self.qoffsets = np.ascontiguousarray(qoffsets, dtype=np.int32)
        self.n = int(self.qoffsets[-1])
        self.stride = max(k, 1)
        self.h = ctypes.c_void_p()
        code = 2 + 2 * tune[0] + int(tune[1]) if isinstance(tune, tuple) else int(bool(tune))
        check(lib().qr_batch_plan_create(batch.h, self.points.ptr, ptr(self.qoffsets), k, code,
                                         ctypes.byref(self.h)))

    def run(self, graph: bool = True) -> float:
        ms = ctypes.c_double()
        fn = lib().qr_batch_plan_run if graph else lib().qr_batch_plan_run_eager
        check(fn(self.h, ctypes.byref(ms)))
        return ms.value

    def read(self) -> np.ndarray:
        out = np.zeros((self.n, self.stride, 2), np.uint32)
        check(lib().qr_batch_plan_read(self.h, ptr(out)))
        return out

    def config(self):
        v, m, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(lib().qr_batch_plan_config(self.h, ctypes.byref(v), ctypes.byref(m), ctypes.byref(t)))
        return {"variant": BATCH_VARIANTS[v.value], "morton": bool(m.value),
                "tuned_ms": t.value}

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            lib().qr_batch_plan_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
