"""The CPU LBVH restatement (node bytes the GPU build must reproduce).

Checked structurally: every internal node covers a contiguous range of leaf
slots whose keys share the common prefix, its split is where the next bit
flips, parent boxes are the bitwise compare-select merge of the children,
leaves hold their primitive's box, and the permutation sorts the Morton keys.
"""
import numpy as np


def morton_keys(cen, lo, hi):
    ext = (hi - lo).astype(np.float32)
    q = ((cen - lo) / ext).astype(np.float32) * np.float32(1023.0)
    q = (q + np.float32(0.5)).astype(np.float32)
    q = np.clip(q, 0, 1023).astype(np.int64)
    code = np.zeros(cen.shape[0], np.int64)
    for b in range(10):
        for a in range(3):
            code |= ((q[:, a] >> b) & 1) << (3 * b + a)
    return (code << 32) | np.arange(cen.shape[0])


def test_lbvh_structure(orc, scenes):
    tris = scenes.terrain(24, 3)
    n = tris.shape[0]
    raw, ids, depth = orc.lbvh_nodes(0, tris)
    nodes = raw.view(np.float32).reshape(-1, 8)
    links = raw.view(np.int32).reshape(-1, 8)[:, 6:8]
    t = tris.reshape(-1, 3, 3)
    cen = ((t[:, 0] + t[:, 1]) + t[:, 2]) / np.float32(3.0)
    keys = morton_keys(cen, cen.min(0), cen.max(0))
    assert np.array_equal(ids, np.argsort(keys, kind="stable"))
    sk = keys[ids]
    # leaves
    for k in range(n):
        assert links[n - 1 + k, 0] == -(k + 1) and links[n - 1 + k, 1] == 1
        tb = t[ids[k]]
        assert np.array_equal(nodes[n - 1 + k, :3], tb.min(0))
        assert np.array_equal(nodes[n - 1 + k, 3:6], tb.max(0))
    # internal: ranges, splits, boxes
    rng = {}

    def cover(i):
        if i >= n - 1:
            return (i - (n - 1), i - (n - 1))
        if i not in rng:
            a, _ = cover(links[i, 0])
            _, b = cover(links[i, 1])
            rng[i] = (a, b)
        return rng[i]

    seen = np.zeros(2 * n - 1, np.int64)
    for i in range(n - 1):
        l, r = links[i]
        seen[l] += 1
        seen[r] += 1
        a, b = cover(i)
        la, lb = cover(l)
        ra, rb = cover(r)
        assert la == a and rb == b and lb + 1 == ra
        top = 63 - int(sk[a] ^ sk[b]).bit_length() + 1  # common prefix length
        bit = 63 - top
        assert (int(sk[lb]) >> bit) & 1 == 0 and (int(sk[ra]) >> bit) & 1 == 1
        L, R = nodes[l], nodes[r]
        want_lo = np.where(L[:3] < R[:3], L[:3], R[:3])
        want_hi = np.where(L[3:6] > R[3:6], L[3:6], R[3:6])
        assert np.array_equal(nodes[i, :3].view(np.uint32), want_lo.view(np.uint32))
        assert np.array_equal(nodes[i, 3:6].view(np.uint32), want_hi.view(np.uint32))
    assert seen[0] == 0 and np.all(seen[1:] == 1)
    assert 10 <= depth <= 64
