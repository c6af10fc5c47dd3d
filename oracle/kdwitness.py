"""Independent CPU witness for point queries: scipy's cKDTree.

TEST INFRASTRUCTURE ONLY (imported by tests/ and bench.py's baseline leg).  The
reference's own external check (bench/scipy_kd_bench.py) runs cKDTree on the same
FP32 inputs (a QRCD dump) and re-ranks its answers by the FP32 squared distance
((dx*dx + dy*dy) + dz*dz with d = q - p) and the point index, so that its
fingerprints can equal the GPU's.  This module does the same with a guard that
makes the witness exact rather than approximate:

* kNN: cKDTree returns the k + extra nearest points in FP64; they are re-ranked by
  (FP32 dist2, index) and the first k kept.  A point outside that candidate set
  lies at FP64 distance >= the last candidate's, so it cannot enter the FP32 top
  k when the k-th FP32 dist2 is below the last candidate's squared distance by more
  than the FP32 rounding of a squared distance (a relative 1e-5 here); the function
  asserts that margin for every query.
* radius: cKDTree returns every point within r (1 + 1e-5) in FP64 -- a superset of
  the FP32 answer -- and the FP32 rule dist2 <= r*r (both in FP32) selects from it.
"""
from __future__ import annotations

import numpy as np

MARGIN = 1e-5


def dist2_f32(q: np.ndarray, p: np.ndarray) -> np.ndarray:
    """FP32 squared distances of q (3,) to p (m, 3), summed x, y then z."""
    d = np.asarray(q, np.float32)[None, :] - np.asarray(p, np.float32)
    return (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]


def _tree(points, leafsize: int = 16):
    from scipy.spatial import cKDTree

    return cKDTree(np.asarray(points, np.float64), leafsize=leafsize)


def knn(points, queries, k: int, extra: int = 8, tree=None):
    """(dist2 float32 (m, k), ids int32 (m, k)) in (dist2, index) order; rows with
    fewer than k points are padded with (inf, -1)."""
    pts = np.asarray(points, np.float32)
    qs = np.asarray(queries, np.float32)
    tree = tree if tree is not None else _tree(pts)
    kk = min(k + extra, pts.shape[0])
    dist, idx = tree.query(qs.astype(np.float64), k=kk)
    dist = np.asarray(dist).reshape(len(qs), kk)
    idx = np.asarray(idx).reshape(len(qs), kk)
    out_d = np.full((len(qs), k), np.inf, np.float32)
    out_i = np.full((len(qs), k), -1, np.int32)
    for i in range(len(qs)):
        ids = idx[i]
        d2 = dist2_f32(qs[i], pts[ids])
        order = np.lexsort((ids, d2))[:k]
        out_d[i, :len(order)] = d2[order]
        out_i[i, :len(order)] = ids[order]
        if kk < pts.shape[0] and len(order) == k:
            last = float(dist[i, -1]) ** 2
            assert float(out_d[i, k - 1]) < last * (1.0 - MARGIN), (
                f"query {i}: ties reach the candidate horizon; raise extra")
    return out_d, out_i


def radius(points, queries, r: float, tree=None):
    """(offsets int64 (m + 1), dist2 float32, ids int32) in (dist2, index) order per
    query, for dist2 <= r*r in FP32."""
    pts = np.asarray(points, np.float32)
    qs = np.asarray(queries, np.float32)
    tree = tree if tree is not None else _tree(pts)
    r32 = np.float32(r)
    r2 = r32 * r32
    lists = tree.query_ball_point(qs.astype(np.float64), r=float(r32) * (1.0 + MARGIN))
    offs = np.zeros(len(qs) + 1, np.int64)
    d_parts, i_parts = [], []
    for i, lst in enumerate(lists):
        ids = np.asarray(lst, np.int64)
        d2 = dist2_f32(qs[i], pts[ids]) if ids.size else np.zeros(0, np.float32)
        keep = d2 <= r2
        ids, d2 = ids[keep], d2[keep]
        order = np.lexsort((ids, d2))
        d_parts.append(d2[order])
        i_parts.append(ids[order].astype(np.int32))
        offs[i + 1] = offs[i] + ids.size
    d = np.concatenate(d_parts) if d_parts else np.zeros(0, np.float32)
    ii = np.concatenate(i_parts) if i_parts else np.zeros(0, np.int32)
    return offs, d.astype(np.float32), ii
