"""Differentiable queries: torch.autograd.Function wrappers over the adjoint kernels.

The forward pass is the plain GPU path (same kernels, bit-identical dist2 and t)
plus a forward-with-save companion: the primitive index of every result and, for
triangles, the closest-feature region (for Gaussians the Eberly root).  The
backward pass hands those to the hand-written adjoint kernels of
libquarry_hip.so, which read the one saved primitive per query and never
traverse the tree again.

Differentiable quantities and their parameters:
  * closest_point_dist2(points, vertices, faces) -> dist2 w.r.t. points, vertices
  * ray_hit_t(origins, dirs, vertices, faces)   -> hit t w.r.t. origins, dirs, vertices
  * point_closest_dist2(points, cloud)          -> dist2 w.r.t. points, cloud
  * gaussian_closest_dist2(points, means, rots, scales)
                                                -> dist2 w.r.t. points, means, rots, scales
Every function also returns the (non-differentiable) index of the primitive the
value belongs to (-1: none; ray misses have t = inf and zero gradient).

PyTorch supplies device memory and the stream: every call runs on torch's current
stream.  Inputs must be float32 CUDA tensors.  torch must be imported before the
library is first loaded, so that both use the same HIP runtime; importing this
module does that.
"""
from __future__ import annotations

import ctypes

import torch

from . import _abi as A

_ORDER = {"input": A.ORDER_INPUT, "morton": A.ORDER_MORTON}


def _p(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _on_stream() -> None:
    A.check(A.lib().qr_set_stream(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))


def _f32(t: torch.Tensor, cols: int, what: str) -> torch.Tensor:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{what} must be a CUDA tensor")
    if t.dtype != torch.float32:
        raise ValueError(f"{what} must be float32")
    if t.dim() != 2 or t.shape[1] != cols:
        raise ValueError(f"{what} must have shape (N, {cols})")
    return t.contiguous()


def _build(kind: int, items: torch.Tensor, wide: bool, builder: int) -> A.Bvh:
    _on_stream()
    bvh = A.Bvh.from_device(kind, items.data_ptr(), items.shape[0], builder)
    if wide:
        bvh.build_wide()
    return bvh


class _ClosestDist2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, items, kind, order, wide, builder):
        n = points.shape[0]
        bvh = _build(kind, items.detach(), wide, builder)
        res = torch.empty((n, 2), dtype=torch.int32, device=points.device)
        aux = torch.empty(n, dtype=torch.int32, device=points.device)
        ms = ctypes.c_double()
        _on_stream()
        A.check(A.lib().qr_closest_device(bvh.h, _p(points), n, _p(res), order, int(wide), ctypes.byref(ms)))
        A.check(A.lib().qr_closest_save_aux(bvh.h, _p(points), n, _p(res), _p(aux)))
        ctx.bvh = bvh
        ctx.save_for_backward(points, items, res, aux)
        dist2 = res[:, 0].view(torch.float32).clone()
        index = res[:, 1].to(torch.int64)
        ctx.mark_non_differentiable(index)
        return dist2, index

    @staticmethod
    def backward(ctx, g_dist2, g_index):
        points, items, res, aux = ctx.saved_tensors
        n = points.shape[0]
        g = g_dist2.contiguous().to(torch.float32)
        gp = torch.empty_like(points) if ctx.needs_input_grad[0] else None
        gi = torch.zeros_like(items) if ctx.needs_input_grad[1] else None
        _on_stream()
        A.check(A.lib().qr_adjoint_closest(ctx.bvh.h, _p(points), n, _p(res), _p(aux), _p(g), _p(gp), _p(gi)))
        return gp, gi, None, None, None, None


class _RayHitT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, origins, dirs, tris, tmax, order, wide, builder):
        n = origins.shape[0]
        bvh = _build(A.PRIM_TRI, tris.detach(), wide, builder)
        rays = torch.empty((n, 8), dtype=torch.float32, device=origins.device)
        rays[:, 0:3] = origins.detach()
        rays[:, 3:6] = dirs.detach()
        rays[:, 6] = 1e-6
        rays[:, 7] = tmax
        res = torch.empty(n, dtype=torch.int64, device=origins.device)
        hit = torch.empty(n, dtype=torch.int32, device=origins.device)
        ms = ctypes.c_double()
        _on_stream()
        A.check(A.lib().qr_raycast_save_device(bvh.h, _p(rays), n, _p(res), _p(hit), order, int(wide),
                                               ctypes.byref(ms)))
        ctx.bvh = bvh
        ctx.save_for_backward(rays, tris, res, hit)
        hit_mask = (res >> 32) != 0
        t = (res & 0xFFFFFFFF).to(torch.int32).view(torch.float32).clone()
        t = torch.where(hit_mask, t, torch.full_like(t, float("inf")))
        index = hit.to(torch.int64)
        ctx.mark_non_differentiable(index)
        return t, index

    @staticmethod
    def backward(ctx, g_t, g_index):
        rays, tris, res, hit = ctx.saved_tensors
        n = rays.shape[0]
        g = torch.nan_to_num(g_t.contiguous().to(torch.float32), nan=0.0, posinf=0.0, neginf=0.0)
        go = torch.empty((n, 3), dtype=torch.float32, device=rays.device) if ctx.needs_input_grad[0] else None
        gd = torch.empty((n, 3), dtype=torch.float32, device=rays.device) if ctx.needs_input_grad[1] else None
        gt = torch.zeros_like(tris) if ctx.needs_input_grad[2] else None
        _on_stream()
        A.check(A.lib().qr_adjoint_ray(ctx.bvh.h, _p(rays), n, _p(res), _p(hit), _p(g), _p(go), _p(gd), _p(gt)))
        return go, gd, gt, None, None, None, None


def _tris(vertices: torch.Tensor, faces: torch.Tensor) -> torch.Tensor:
    v = _f32(vertices, 3, "vertices")
    if faces.dim() != 2 or faces.shape[1] != 3:
        raise ValueError("faces must have shape (F, 3)")
    return v[faces.long()].reshape(-1, 9).contiguous()


def closest_point_dist2(points, vertices, faces, order: str = "morton", wide: bool = True,
                        builder: str = "lbvh"):
    """Squared distance from each point to the mesh and the closest face index."""
    p = _f32(points, 3, "points")
    return _ClosestDist2.apply(p, _tris(vertices, faces), A.PRIM_TRI, _ORDER[order], wide,
                               A.BUILDER_SAH if builder == "sah" else A.BUILDER_LBVH)


def ray_hit_t(origins, dirs, vertices, faces, tmax: float = float("inf"), order: str = "input",
              wide: bool = True, builder: str = "lbvh"):
    """Closest-hit t of each ray (inf on a miss) and the hit face index (-1)."""
    o = _f32(origins, 3, "origins")
    d = _f32(dirs, 3, "dirs")
    if o.shape[0] != d.shape[0]:
        raise ValueError("origins and dirs must match")
    orders = {"input": A.ORDER_INPUT, "morton": A.ORDER_MORTON, "morton_direction": A.ORDER_MORTON_DIRECTION}
    return _RayHitT.apply(o, d, _tris(vertices, faces), float(tmax), orders[order], wide,
                          A.BUILDER_SAH if builder == "sah" else A.BUILDER_LBVH)


def point_closest_dist2(points, cloud, order: str = "morton", wide: bool = True):
    """Squared distance from each query to the nearest cloud point, and its index."""
    return _ClosestDist2.apply(_f32(points, 3, "points"), _f32(cloud, 3, "cloud"), A.PRIM_POINT,
                               _ORDER[order], wide, A.BUILDER_LBVH)


def gaussian_closest_dist2(points, means, rots, scales, order: str = "morton", wide: bool = True):
    """Squared distance from each query to the nearest solid ellipsoid (0 inside)."""
    items = torch.cat([_f32(means, 3, "means"), _f32(rots, 4, "rots"), _f32(scales, 3, "scales")], dim=1)
    return _ClosestDist2.apply(_f32(points, 3, "points"), items.contiguous(), A.PRIM_GAUSS,
                               _ORDER[order], wide, A.BUILDER_LBVH)
