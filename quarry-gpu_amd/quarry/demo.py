"""Demo scenes of the Python module: demo_mesh() and demo_gaussians().

Restates the reference's scene generators (src/core/scene.cpp:8-153,
src/core/gaussian.cpp:11-55) over its PCG32 stream (src/core/rng.h:7-38): the
integer stream and the float arithmetic follow the C++ code in float32, while
sin/cos/sqrt come from numpy, whose float32 results can differ from the C
library's in the last place -- so a demo scene equals the C++ one up to such
last-place differences, not bit for bit.  These generators only make demo
inputs; no result of the library depends on them.
"""
from __future__ import annotations

import numpy as np

F = np.float32
_MASK = (1 << 64) - 1


class Pcg32:
    """PCG32 (XSH-RR) with the reference's seeding (src/core/rng.h:9-17)."""

    def __init__(self, seed: int = 0, stream: int = 0):
        self.state = 0
        self.inc = ((stream << 1) | 1) & _MASK
        self.next_u32()
        self.state = (self.state + 0x853C49E6748FEA9B) & _MASK
        self.next_u32()
        self.state = (self.state + seed) & _MASK
        self.next_u32()

    def next_u32(self) -> int:
        old = self.state
        self.state = (old * 6364136223846793005 + self.inc) & _MASK
        xorshifted = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xorshifted >> rot) | (xorshifted << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform01(self) -> np.float32:
        return F(self.next_u32() >> 8) * F(1.0 / 16777216.0)

    def uniform(self, lo, hi) -> np.float32:
        lo, hi = F(lo), F(hi)
        return lo + (hi - lo) * self.uniform01()


def terrain_height(x, z):
    x, z = F(x), F(z)
    s, c = np.sin, np.cos
    return (F(3.0) * s(F(0.13) * x) * c(F(0.11) * z)
            + F(1.2) * s(F(0.37) * x + F(0.7)) * s(F(0.31) * z + F(1.3))
            + F(0.4) * s(F(0.90) * x) * s(F(0.80) * z))


def _terrain(grid: int, half: float = 50.0) -> np.ndarray:
    half = F(half)
    step = (F(2.0) * half) / F(grid)
    i = np.arange(grid, dtype=np.float32)
    x0 = -half + step * i[None, :]
    z0 = -half + step * i[:, None]
    x0, z0 = np.broadcast_arrays(x0, z0)
    x1, z1 = x0 + step, z0 + step

    def v(x, z):
        return np.stack([x, terrain_height(x, z), z], axis=-1)

    p00, p10, p11, p01 = v(x0, z0), v(x1, z0), v(x1, z1), v(x0, z1)
    t0 = np.stack([p00, p10, p11], axis=-2)
    t1 = np.stack([p00, p11, p01], axis=-2)
    return np.stack([t0, t1], axis=2).reshape(-1, 3, 3).astype(np.float32)


def _random_tris(count: int, seed: int) -> np.ndarray:
    rng = Pcg32(seed, 1)
    out = []
    while len(out) < count:
        a = np.array([rng.uniform(-50, 50) for _ in range(3)], np.float32)
        b = a + np.array([rng.uniform(-5, 5) for _ in range(3)], np.float32)
        c = a + np.array([rng.uniform(-5, 5) for _ in range(3)], np.float32)
        if np.linalg.norm(np.cross(b - a, c - a).astype(np.float32)) < F(1e-4):
            continue
        out.append(np.stack([a, b, c]))
    return np.array(out, np.float32).reshape(-1, 3, 3)


def _box(cx, y0, cz, w, h, d) -> np.ndarray:
    x0, x1 = cx - F(0.5) * w, cx + F(0.5) * w
    z0, z1 = cz - F(0.5) * d, cz + F(0.5) * d
    y1 = y0 + h
    b00, b10, b11, b01 = (x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)
    t00, t10, t11, t01 = (x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1)
    faces = [(b00, t00, t10), (b00, t10, b10), (b10, t10, t11), (b10, t11, b11),
             (b11, t11, t01), (b11, t01, b01), (b01, t01, t00), (b01, t00, b00),
             (t00, t01, t11), (t00, t11, t10)]
    return np.array(faces, np.float32)


def _city(buildings: int, seed: int) -> np.ndarray:
    if buildings <= 0:
        return _terrain(128)
    side = max(1, int(np.ceil(np.sqrt(float(buildings)))))
    spacing = F(8.0)
    half = F(0.5) * spacing * F(side)
    parts = [_terrain(128, half)]
    rng = Pcg32(seed, 4)
    for i in range(buildings):
        bx, bz = i % side, i // side
        cx = -half + (F(bx) + F(0.5)) * spacing + rng.uniform(-1.2, 1.2)
        cz = -half + (F(bz) + F(0.5)) * spacing + rng.uniform(-1.2, 1.2)
        w = rng.uniform(1.5, 5.0)
        d = rng.uniform(1.5, 5.0)
        u = rng.uniform01()
        h = F(4.0) + F(56.0) * u * u * u
        parts.append(_box(cx, terrain_height(cx, cz) - F(0.6), cz, w, h, d))
    return np.concatenate(parts).astype(np.float32)


def demo_mesh(kind: str = "terrain", grid: int = 64, random_tris: int = 0, seed: int = 1,
              buildings: int = 65536):
    """(vertices (3M, 3) float32, indices (M, 3) int32) of a demo scene
    (quarry_py.cpp:858-883); kind is terrain | random | mixed | city."""
    if kind == "terrain":
        tris = _terrain(grid)
    elif kind == "random":
        tris = _random_tris(random_tris, seed)
    elif kind == "mixed":
        tris = np.concatenate([_terrain(grid), _random_tris(random_tris, seed)])
    elif kind == "city":
        tris = _city(buildings, seed)
    else:
        tris = np.zeros((0, 3, 3), np.float32)
    m = tris.shape[0]
    return tris.reshape(-1, 3).copy(), np.arange(3 * m, dtype=np.int32).reshape(m, 3)


def demo_gaussians(kind: str = "sheet", count: int = 131072, seed: int = 1):
    """(means (N, 3), rot (N, 4) xyzw, scales (N, 3)) of a demo Gaussian cloud
    (quarry_py.cpp:826-856); kind is sheet (means on the terrain) or volume."""
    rng = Pcg32(seed, 5)
    half = F(50.0)
    means = np.zeros((max(count, 0), 3), np.float32)
    rot = np.zeros((max(count, 0), 4), np.float32)
    scale = np.zeros((max(count, 0), 3), np.float32)
    two_pi = F(6.28318530717958647692)
    for i in range(max(count, 0)):
        if kind == "volume":
            means[i] = [rng.uniform(-half, half) for _ in range(3)]
        else:
            x = rng.uniform(-half, half)
            z = rng.uniform(-half, half)
            means[i] = [x, terrain_height(x, z), z]
        u1, u2, u3 = rng.uniform01(), rng.uniform01(), rng.uniform01()
        s1, s2 = np.sqrt(F(1.0) - u1), np.sqrt(u1)
        rot[i] = [s1 * np.sin(two_pi * u2), s1 * np.cos(two_pi * u2),
                  s2 * np.sin(two_pi * u3), s2 * np.cos(two_pi * u3)]
        major = rng.uniform(0.08, 0.50)
        scale[i] = [major, major * rng.uniform(0.40, 1.0), major * rng.uniform(0.40, 1.0)]
    return means, rot, scale
