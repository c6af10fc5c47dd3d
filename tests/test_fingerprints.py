"""Fingerprint recipes (src/core/fingerprint.h:16-94) recomputed independently."""
import numpy as np

OFF, PRIME, MASK = 0xCBF29CE484222325, 0x100000001B3, (1 << 64) - 1


def fnv(values):
    h = OFF
    for v in values:
        h ^= int(v) & MASK
        h = (h * PRIME) & MASK
    return h


def test_ray_hash(orc):
    rng = np.random.default_rng(1)
    t = rng.uniform(0, 10, 500).astype(np.float32)
    hit = rng.random(500) < 0.7
    v = np.where(hit, (np.uint64(1) << np.uint64(32)) | t.view(np.uint32).astype(np.uint64), np.uint64(0))
    h, hits, s = orc.ray_hash(v)
    assert h == fnv(v.tolist())
    assert hits == int(hit.sum())
    assert abs(s - float(t[hit].astype(np.float64).sum())) < 1e-6
    assert orc.ray_hash(np.zeros(0, np.uint64))[0] == OFF


def test_closest_hash(orc):
    rec = np.array([[[0x3F800000, 7]], [[0, 0xFFFFFFFF]], [[0x40000000, 0]]], np.uint32)
    vals = [(1 << 32) | 0x3F800000, 8, 0, 0, (1 << 32) | 0x40000000, 1]
    h, s = orc.closest_hash(rec)
    assert h == fnv(vals) and s == 3.0


def test_list_hashes(orc):
    rec = np.array([[[0x3F800000, 3], [0x40000000, 1], [0, 0xFFFFFFFF]],
                    [[0, 0xFFFFFFFF], [0, 0xFFFFFFFF], [0, 0xFFFFFFFF]]], np.uint32)
    vals = [2, 0x3F800000, 4, 0x40000000, 2, 0]
    assert orc.knn_hash(rec, 3)[0] == fnv(vals)
    offs = np.array([0, 2, 2], np.int32)
    keys = np.array([(0x3F800000 << 32) | 3, (0x40000000 << 32) | 1], np.uint64)
    assert orc.csr_hash(offs, keys)[0] == fnv(vals)
    # the length is mixed first: an empty list differs from a missing entry
    assert orc.csr_hash(np.array([0, 0], np.int32), np.zeros(0, np.uint64))[0] == fnv([0])
