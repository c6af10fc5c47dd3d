"""Seeded synthetic scenes for the benchmark workloads (FP32 arrays, numpy).

Each generator returns contiguous float32 arrays in the library's wire layouts:
triangles (n, 9), points (n, 3), Gaussians (n, 10: mean, quat xyzw, scale),
rays (n, 8: origin, dir, tMin, tMax).  Randomness comes from numpy's PCG64
seeded per array, so a (generator, seed) pair always produces the same bytes on
one machine; tests and benchmarks feed the same arrays to the GPU and to the
CPU checker in one process.

Choices the workload descriptions leave open are fixed here and repeated in
SPEC below:
  terrain   z-up heightfield, unit xy spacing, (g+1)^2 vertices, two triangles
            per cell split as (p00, p10, p11), (p00, p11, p01); height = sum of
            8 sinusoids sum_i (8/8) sin(k_i . (x, y) + phi_i) with wavelengths
            log-uniform in [16, 128] cells and uniform directions, plus U(+-0.25)
  city      side = ceil(sqrt(B)) columns of an 8-unit jittered grid (+-1.2),
            footprint U[0.3, 0.9] x 8 per axis, height lognormal(1.0, 0.8),
            10 triangles per building (4 walls + roof), ground = 128^2 terrain
            cells spanning the city (2 * 128^2 triangles)
  points    half on a sphere shell of radius 16 centred in the 64-cube, half
            uniform in the cube [0, 64]^3
  gaussians means uniform in [0, 64]^3, per-axis scale log-uniform [0.05, 1],
            Shoemake uniform quaternions
  icosphere radius 10, 7 subdivisions (327,680 triangles), per-vertex N(0, 0.1)
  batch     S scenes, n_s = floor(exp(U[ln 256, ln 65536])) triangles; even s: a
            terrain patch of ceil(sqrt(n_s/2))^2 cells (first n_s triangles kept),
            odd s: a city block with an 8x8-cell ground and (n_s - 128) // 10
            buildings; q_s = floor(exp(U[ln 64, ln 8192])) queries uniform in the
            scene's box, the first ceil(q_s/2) for closest, the rest for kNN k=8
"""
from __future__ import annotations

import math

import numpy as np

SPEC = __doc__

TRI_F, PT_F, GS_F, RAY_F = 9, 3, 10, 8


def _rng(seed: int, stream: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64([seed, stream]))


# ---- terrain ------------------------------------------------------------------------

def terrain_heights(cells: int, seed: int, amp: float = 8.0, noise: float = 0.25,
                    x0: float = 0.0, y0: float = 0.0) -> np.ndarray:
    """(cells+1, cells+1) float32 heights at integer grid points (row = y)."""
    rng = _rng(seed, 1)
    waves = 8
    wl = np.exp(rng.uniform(math.log(16.0), math.log(128.0), waves))
    ang = rng.uniform(0.0, 2.0 * math.pi, waves)
    phase = rng.uniform(0.0, 2.0 * math.pi, waves)
    kx = 2.0 * math.pi / wl * np.cos(ang)
    ky = 2.0 * math.pi / wl * np.sin(ang)
    xs = x0 + np.arange(cells + 1, dtype=np.float64)
    ys = y0 + np.arange(cells + 1, dtype=np.float64)
    X, Y = np.meshgrid(xs, ys)
    h = np.zeros_like(X)
    for i in range(waves):
        h += (amp / waves) * np.sin(kx[i] * X + ky[i] * Y + phase[i])
    h += rng.uniform(-noise, noise, X.shape)
    return h.astype(np.float32)


def grid_triangles(xs: np.ndarray, ys: np.ndarray, h: np.ndarray) -> np.ndarray:
    """Two triangles per cell of a z-up heightfield; xs/ys are the vertex coordinates."""
    X, Y = np.meshgrid(xs.astype(np.float32), ys.astype(np.float32))
    V = np.stack([X, Y, h.astype(np.float32)], axis=-1)  # (ny+1, nx+1, 3)
    p00 = V[:-1, :-1]
    p10 = V[:-1, 1:]
    p11 = V[1:, 1:]
    p01 = V[1:, :-1]
    t0 = np.concatenate([p00, p10, p11], axis=-1)
    t1 = np.concatenate([p00, p11, p01], axis=-1)
    tris = np.stack([t0, t1], axis=2)  # (ny, nx, 2, 9)
    return np.ascontiguousarray(tris.reshape(-1, TRI_F), dtype=np.float32)


def terrain(cells: int, seed: int = 1) -> np.ndarray:
    """cells^2 * 2 triangles at unit spacing over [0, cells]^2."""
    h = terrain_heights(cells, seed)
    idx = np.arange(cells + 1, dtype=np.float32)
    return grid_triangles(idx, idx, h)


def terrain_vertex_ids(cells: int) -> np.ndarray:
    """(ntris, 3) vertex ids (row-major over the (cells+1)^2 grid) of terrain()."""
    n1 = cells + 1
    j, i = np.meshgrid(np.arange(cells), np.arange(cells))  # i row (y), j col (x)
    v00 = i * n1 + j
    v10 = v00 + 1
    v11 = v00 + n1 + 1
    v01 = v00 + n1
    a = np.stack([v00, v10, v11], axis=-1)
    b = np.stack([v00, v11, v01], axis=-1)
    return np.stack([a, b], axis=2).reshape(-1, 3).astype(np.int64)


def bounds_of(tris: np.ndarray) -> np.ndarray:
    t = tris.reshape(-1, 3, 3)
    return np.concatenate([t.min(axis=(0, 1)), t.max(axis=(0, 1))]).astype(np.float32)


def down_rays(tris: np.ndarray, count: int, seed: int = 2) -> np.ndarray:
    """Rays straight down, origins uniform over the xy extent at z = max height + 1."""
    b = bounds_of(tris)
    rng = _rng(seed, 2)
    r = np.zeros((count, RAY_F), dtype=np.float32)
    r[:, 0] = rng.uniform(b[0], b[3], count)
    r[:, 1] = rng.uniform(b[1], b[4], count)
    r[:, 2] = b[5] + np.float32(1.0)
    r[:, 5] = -1.0
    r[:, 6] = np.float32(1e-6)
    r[:, 7] = np.inf
    return r


def query_points(lo: np.ndarray, hi: np.ndarray, count: int, seed: int = 3) -> np.ndarray:
    rng = _rng(seed, 3)
    return np.ascontiguousarray(rng.uniform(lo, hi, (count, 3)), dtype=np.float32)


def displace_vertices(cells: int, tris: np.ndarray, frac: float = 0.01, sigma: float = 0.05,
                      seed: int = 4):
    """Moves a random `frac` of the terrain's grid vertices by N(0, sigma) per axis.

    Returns (new triangles, dirty triangle indices)."""
    rng = _rng(seed, 4)
    nv = (cells + 1) ** 2
    m = int(round(frac * nv))
    chosen = rng.choice(nv, size=m, replace=False)
    delta = np.zeros((nv, 3), dtype=np.float32)
    delta[chosen] = rng.normal(0.0, sigma, (m, 3)).astype(np.float32)
    vid = terrain_vertex_ids(cells)
    t = tris.reshape(-1, 3, 3).copy()
    t += delta[vid]
    moved = np.zeros(nv, dtype=bool)
    moved[chosen] = True
    dirty = np.nonzero(moved[vid].any(axis=1))[0].astype(np.int32)
    return np.ascontiguousarray(t.reshape(-1, TRI_F)), dirty


# ---- city -----------------------------------------------------------------------------

def _box_triangles(cx, cy, w, d, z0, h) -> np.ndarray:
    """10 triangles per building: 4 walls + roof (no floor), z-up."""
    x0, x1 = cx - 0.5 * w, cx + 0.5 * w
    y0, y1 = cy - 0.5 * d, cy + 0.5 * d
    z1 = z0 + h

    def P(x, y, z):
        return np.stack([x, y, z], axis=-1)

    b00, b10, b11, b01 = P(x0, y0, z0), P(x1, y0, z0), P(x1, y1, z0), P(x0, y1, z0)
    t00, t10, t11, t01 = P(x0, y0, z1), P(x1, y0, z1), P(x1, y1, z1), P(x0, y1, z1)
    tris = [
        (b00, t00, t10), (b00, t10, b10), (b10, t10, t11), (b10, t11, b11),
        (b11, t11, t01), (b11, t01, b01), (b01, t01, t00), (b01, t00, b00),
        (t00, t01, t11), (t00, t11, t10),
    ]
    out = np.stack([np.concatenate(t, axis=-1) for t in tris], axis=1)  # (B, 10, 9)
    return out.reshape(-1, TRI_F)


def city(buildings: int, seed: int = 1, spacing: float = 8.0, ground_cells: int = 128):
    rng = _rng(seed, 5)
    side = int(math.ceil(math.sqrt(buildings)))
    half = 0.5 * spacing * side
    i = np.arange(buildings)
    bx, by = i % side, i // side
    cx = (-half + (bx + 0.5) * spacing + rng.uniform(-1.2, 1.2, buildings)).astype(np.float32)
    cy = (-half + (by + 0.5) * spacing + rng.uniform(-1.2, 1.2, buildings)).astype(np.float32)
    w = (rng.uniform(0.3, 0.9, buildings) * spacing).astype(np.float32)
    d = (rng.uniform(0.3, 0.9, buildings) * spacing).astype(np.float32)
    h = rng.lognormal(1.0, 0.8, buildings).astype(np.float32)
    # Ground: a gentle terrain under the city; buildings start 0.6 below it.
    gstep = 2.0 * half / ground_cells
    gx = (-half + gstep * np.arange(ground_cells + 1)).astype(np.float32)
    gh = (0.5 * terrain_heights(ground_cells, seed + 7, amp=2.0, noise=0.0)).astype(np.float32)
    ground = grid_triangles(gx, gx, gh)
    # ground height at each building centre (bilinear on the vertex grid)
    fx = np.clip((cx + half) / gstep, 0, ground_cells - 1e-3)
    fy = np.clip((cy + half) / gstep, 0, ground_cells - 1e-3)
    ix, iy = fx.astype(int), fy.astype(int)
    ax, ay = fx - ix, fy - iy
    z = ((1 - ax) * (1 - ay) * gh[iy, ix] + ax * (1 - ay) * gh[iy, ix + 1] +
         (1 - ax) * ay * gh[iy + 1, ix] + ax * ay * gh[iy + 1, ix + 1]).astype(np.float32)
    blds = _box_triangles(cx, cy, w, d, z - np.float32(0.6), h)
    return np.ascontiguousarray(np.concatenate([ground, blds.astype(np.float32)]), dtype=np.float32)


def cosine_down_rays(tris: np.ndarray, count: int, seed: int = 6) -> np.ndarray:
    """Origins uniform over the xy extent above the tallest point, cosine-weighted
    directions about -z."""
    b = bounds_of(tris)
    rng = _rng(seed, 6)
    r = np.zeros((count, RAY_F), dtype=np.float32)
    r[:, 0] = rng.uniform(b[0], b[3], count)
    r[:, 1] = rng.uniform(b[1], b[4], count)
    r[:, 2] = b[5] + np.float32(1.0)
    u1 = rng.uniform(0.0, 1.0, count)
    u2 = rng.uniform(0.0, 1.0, count)
    rad = np.sqrt(u1)
    phi = 2.0 * math.pi * u2
    r[:, 3] = rad * np.cos(phi)
    r[:, 4] = rad * np.sin(phi)
    r[:, 5] = -np.sqrt(np.maximum(0.0, 1.0 - u1))
    r[:, 6] = np.float32(1e-6)
    r[:, 7] = np.inf
    return r


# ---- points and Gaussians ---------------------------------------------------------------

def point_cloud(count: int, seed: int = 1, side: float = 64.0, shell_radius: float = 16.0) -> np.ndarray:
    rng = _rng(seed, 7)
    ns = count // 2
    v = rng.normal(size=(ns, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    shell = 0.5 * side + shell_radius * v
    cube = rng.uniform(0.0, side, (count - ns, 3))
    return np.ascontiguousarray(np.concatenate([shell, cube]), dtype=np.float32)


def shoemake_quats(rng: np.random.Generator, count: int) -> np.ndarray:
    u1, u2, u3 = (rng.uniform(0.0, 1.0, count) for _ in range(3))
    s1, s2 = np.sqrt(1.0 - u1), np.sqrt(u1)
    return np.stack([s1 * np.sin(2 * math.pi * u2), s1 * np.cos(2 * math.pi * u2),
                     s2 * np.sin(2 * math.pi * u3), s2 * np.cos(2 * math.pi * u3)], axis=1)


def gaussians(count: int, seed: int = 1, side: float = 64.0) -> np.ndarray:
    rng = _rng(seed, 8)
    g = np.zeros((count, GS_F), dtype=np.float32)
    g[:, 0:3] = rng.uniform(0.0, side, (count, 3))
    g[:, 3:7] = shoemake_quats(rng, count)
    g[:, 7:10] = np.exp(rng.uniform(math.log(0.05), math.log(1.0), (count, 3)))
    return g


# ---- icosphere ----------------------------------------------------------------------------

def icosphere(levels: int = 7, radius: float = 10.0, noise: float = 0.1, seed: int = 1):
    """Returns (vertices (V, 3) float32, faces (F, 3) int32, triangles (F, 9))."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
             (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = np.array(verts, dtype=np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array(faces, dtype=np.int64)
    for _ in range(levels):
        a, b, c = F[:, 0], F[:, 1], F[:, 2]
        edges = np.concatenate([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)])
        key = np.sort(edges, axis=1)
        uniq, inv = np.unique(key[:, 0] * (len(V) + 1) + key[:, 1], return_inverse=True)
        e0 = uniq // (len(V) + 1)
        e1 = uniq % (len(V) + 1)
        mid = V[e0] + V[e1]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        mids = len(V) + inv
        m = len(F)
        ab, bc, ca = mids[:m], mids[m:2 * m], mids[2 * m:]
        V = np.concatenate([V, mid])
        F = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)])
    rng = _rng(seed, 9)
    V = V * radius + rng.normal(0.0, noise, V.shape)
    V32 = V.astype(np.float32)
    tris = np.ascontiguousarray(V32[F].reshape(-1, TRI_F))
    return V32, F.astype(np.int32), tris


def sphere_rays(count: int, seed: int = 10, origin_radius: float = 30.0,
                target_radius: float = 5.0) -> np.ndarray:
    rng = _rng(seed, 10)
    o = rng.normal(size=(count, 3))
    o *= origin_radius / np.linalg.norm(o, axis=1, keepdims=True)
    tgt = rng.normal(size=(count, 3))
    tgt /= np.linalg.norm(tgt, axis=1, keepdims=True)
    tgt *= target_radius * rng.uniform(0.0, 1.0, (count, 1)) ** (1.0 / 3.0)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((count, RAY_F), dtype=np.float32)
    r[:, 0:3] = o
    r[:, 3:6] = d
    r[:, 6] = np.float32(1e-6)
    r[:, 7] = np.inf
    return r


# ---- batched scenes -------------------------------------------------------------------------

def batch_scene(s: int, n: int, seed: int) -> np.ndarray:
    """Scene s of the batched workload with about n triangles (n exactly for terrain)."""
    if s % 2 == 0:
        cells = int(math.ceil(math.sqrt(n / 2.0)))
        return np.ascontiguousarray(terrain(cells, seed * 100003 + s)[:n])
    b = max(1, (n - 128) // 10)
    return city(b, seed * 100003 + s, ground_cells=8)


def workload_c5(seed: int = 1, scenes: int = 4096, nmin: int = 256, nmax: int = 65536,
                qmin: int = 64, qmax: int = 8192):
    """Returns (items (N, 9), prim offsets (S+1), closest points, closest offsets,
    kNN points, kNN offsets); every array is scene-major."""
    rng = _rng(seed, 11)
    ns = np.floor(np.exp(rng.uniform(math.log(nmin), math.log(nmax), scenes))).astype(np.int64)
    qs = np.floor(np.exp(rng.uniform(math.log(qmin), math.log(qmax), scenes))).astype(np.int64)
    parts, cq, kq = [], [], []
    for s in range(scenes):
        tris = batch_scene(s, int(ns[s]), seed)
        parts.append(tris)
        b = bounds_of(tris)
        q = query_points(b[:3], b[3:], int(qs[s]), seed * 7919 + s)
        h = (int(qs[s]) + 1) // 2
        cq.append(q[:h])
        kq.append(q[h:])
    poff = np.zeros(scenes + 1, np.int32)
    np.cumsum([p.shape[0] for p in parts], out=poff[1:])
    coff = np.zeros(scenes + 1, np.int32)
    np.cumsum([c.shape[0] for c in cq], out=coff[1:])
    koff = np.zeros(scenes + 1, np.int32)
    np.cumsum([k.shape[0] for k in kq], out=koff[1:])
    return (np.ascontiguousarray(np.concatenate(parts)), poff, np.ascontiguousarray(np.concatenate(cq)), coff,
            np.ascontiguousarray(np.concatenate(kq)), koff)


# ---- the benchmark workloads ----------------------------------------------------------------

def workload_c1(seed: int = 1, cells: int = 256, rays: int = 1_000_000):
    tris = terrain(cells, seed)
    return tris, down_rays(tris, rays, seed + 1)


def workload_c2(seed: int = 1, cells: int = 1024, queries: int = 100_000):
    tris = terrain(cells, seed)
    b = bounds_of(tris)
    q = query_points(b[:3] - 4.0, b[3:] + 4.0, queries, seed + 2)
    return tris, q
