"""Writes tests/golden/*.npz: inputs and outputs of the reference's own CPU code.

The reference's CPU sources (src/core, src/cpu of ycl-berserk/quarry-gpu) are
compiled from /root/reference by `make -C oracle ref` into
oracle/_ref/libquarry_ref.so; this script calls that library on small seeded
inputs and stores inputs + expected outputs.  The fixtures are data only; the
oracle is checked against them in tests/test_oracle_golden.py.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def f32p(a):
    return a.ctypes.data_as(O._f32p)


def ref_csr(R, fn, h, pts, r):
    n = pts.shape[0]
    counts = np.zeros(n, np.int32)
    fn(h, f32p(pts), n, r, counts.ctypes.data_as(O._i32p), None, None, 1)
    offs = np.zeros(n + 1, np.int32)
    np.cumsum(counts, out=offs[1:])
    keys = np.zeros(max(int(offs[-1]), 1), np.uint64)
    fn(h, f32p(pts), n, r, counts.ctypes.data_as(O._i32p), offs.ctypes.data_as(O._i32p),
       keys.ctypes.data_as(O._u64p), 1)
    return offs, keys[: int(offs[-1])]


def ref_knn(R, fn, h, pts, k):
    n = pts.shape[0]
    out = np.zeros((n, max(k, 1), 2), np.uint32)
    fn(h, f32p(pts), n, k, out.ctypes.data_as(O._u32p), 1)
    return out


def main() -> None:
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libquarry_ref.so is missing: run `make -C oracle ref`")

    # -- triangles: the reference's "mixed" scene (terrain 24 + 800 random) -------------
    cap = 2 * 24 * 24 + 800
    tris = np.zeros((cap, 9), np.float32)
    b6 = np.zeros(6, np.float32)
    n = R.ref_scene(3, 24, 800, 7, f32p(tris), cap, f32p(b6))
    tris = tris[:n].copy()
    rays_down = np.zeros((1500, 8), np.float32)
    rays_rand = np.zeros((1500, 8), np.float32)
    R.ref_rays(f32p(b6), 0, 1500, 11, f32p(rays_down))
    R.ref_rays(f32p(b6), 1, 1500, 11, f32p(rays_rand))
    pts = np.zeros((400, 3), np.float32)
    R.ref_query_points(f32p(b6), 400, 13, f32p(pts))
    h = R.ref_bvh_new(f32p(tris), n)
    ray_d = np.zeros(1500, np.uint64)
    ray_r = np.zeros(1500, np.uint64)
    R.ref_bvh_ray(h, f32p(rays_down), 1500, ray_d.ctypes.data_as(O._u64p), 1)
    R.ref_bvh_ray(h, f32p(rays_rand), 1500, ray_r.ctypes.data_as(O._u64p), 1)
    closest = ref_knn(R, R.ref_bvh_knn, h, pts, 0)
    knn = ref_knn(R, R.ref_bvh_knn, h, pts, 6)
    r_off, r_keys = ref_csr(R, R.ref_bvh_radius, h, pts, 3.0)
    R.ref_bvh_free(h)
    np.savez_compressed(os.path.join(OUT, "tri_mixed.npz"), tris=tris, rays_down=rays_down,
                        rays_random=rays_rand, points=pts, ray_down=ray_d, ray_random=ray_r,
                        closest=closest, knn6=knn, radius3_offsets=r_off, radius3_keys=r_keys)

    # -- point cloud (kd-tree) ------------------------------------------------------------
    rng = np.random.default_rng(5)
    cloud = rng.uniform(-20, 20, (3000, 3)).astype(np.float32)
    cloud[1000:1010] = cloud[0]  # duplicates exercise the index tie rule
    q = rng.uniform(-24, 24, (300, 3)).astype(np.float32)
    h = R.ref_kd_new(f32p(cloud), cloud.shape[0])
    p_closest = ref_knn(R, R.ref_kd_knn, h, q, 0)
    p_knn = ref_knn(R, R.ref_kd_knn, h, q, 8)
    p_off, p_keys = ref_csr(R, R.ref_kd_radius, h, q, 2.5)
    R.ref_kd_free(h)
    np.savez_compressed(os.path.join(OUT, "points.npz"), cloud=cloud, points=q, closest=p_closest,
                        knn8=p_knn, radius_offsets=p_off, radius_keys=p_keys)

    # -- Gaussians (ellipsoid BVH) -----------------------------------------------------------
    g = np.zeros((1500, 10), np.float32)
    g[:, 0:3] = rng.uniform(-15, 15, (1500, 3))
    qq = rng.normal(size=(1500, 4))
    qq /= np.linalg.norm(qq, axis=1, keepdims=True)
    g[:, 3:7] = qq * rng.uniform(0.5, 2.0, (1500, 1))  # unnormalised on purpose
    g[:, 7:10] = np.exp(rng.uniform(np.log(0.05), np.log(1.5), (1500, 3)))
    gq = rng.uniform(-18, 18, (300, 3)).astype(np.float32)
    h = R.ref_gauss_new(f32p(g), g.shape[0])
    g_closest = ref_knn(R, R.ref_gauss_knn, h, gq, 0)
    g_knn = ref_knn(R, R.ref_gauss_knn, h, gq, 8)
    g_off, g_keys = ref_csr(R, R.ref_gauss_radius, h, gq, 1.0)
    R.ref_gauss_free(h)
    dist = np.array([R.ref_gauss_dist2(f32p(gq[i % 300].copy()), f32p(g[i].copy()))
                     for i in range(1500)], np.float32)
    np.savez_compressed(os.path.join(OUT, "gaussians.npz"), gaussians=g, points=gq,
                        closest=g_closest, knn8=g_knn, radius_offsets=g_off, radius_keys=g_keys,
                        pair_dist2=dist)
    print("wrote", sorted(f for f in os.listdir(OUT) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
