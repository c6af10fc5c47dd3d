"""The CPU LBVH restatement (node bytes the GPU build must reproduce).

Checked structurally: every internal node covers a contiguous range of leaf
slots whose keys share the common prefix, its split is where the next bit
flips, parent boxes are the bitwise compare-select merge of the children,
leaves hold their primitive's box, and the permutation sorts the Morton keys.
"""
import numpy as np
import pytest


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


def _links_depth(raw, n):
    links = raw.view(np.int32).reshape(-1, 8)[:, 6:8]
    return links[: n - 1, 0], links[: n - 1, 1]


def _karras_rule(left, right, n):
    """Karras 2012's numbering: the root is 0 and covers every slot; an internal
    left child is numbered by its range's last slot, an internal right child by its
    range's first slot; leaves are n - 1 + slot."""
    ranges = {0: (0, n - 1)}
    stack = [0]
    while stack:
        i = stack.pop()
        a, b = ranges[i]
        l, r = int(left[i]), int(right[i])
        # the children's ranges follow from the split: left [a, s], right [s + 1, b]
        s = l - (n - 1) if l >= n - 1 else l       # left id is its range's last slot
        assert a <= s < b
        rs = r - (n - 1) if r >= n - 1 else r      # right id is its range's first slot
        assert rs == s + 1, (i, a, b, l, r)
        if l >= n - 1:
            assert s == a                          # a leaf child covers one slot
        else:
            ranges[l] = (a, s)
            stack.append(l)
        if r >= n - 1:
            assert rs == b
        else:
            ranges[r] = (rs, b)
            stack.append(r)
    assert len(ranges) == n - 1                    # every internal node reached once


def _items(kind, pts, rng):
    if kind == 0:
        return np.concatenate([pts, pts, pts], 1).astype(np.float32)
    if kind == 1:
        return pts.astype(np.float32)
    g = np.concatenate([pts, rng.normal(size=(pts.shape[0], 4)),
                        np.exp(rng.uniform(-3, 0, (pts.shape[0], 3)))], 1)
    return g.astype(np.float32)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_karras_numbering_all_kinds(orc, scenes, kind):
    rng = np.random.default_rng(90 + kind)
    pts = rng.uniform(-20, 20, (3000, 3))
    pts[500:700] = pts[3]                          # equal codes, ordered by index
    items = _items(kind, pts, rng)
    raw, ids, depth = orc.lbvh_nodes(kind, items)
    n = items.shape[0]
    left, right = _links_depth(raw, n)
    _karras_rule(left, right, n)


def _dyadic_points(rng, n):
    """Points whose coordinates have few significant bits, so that the triangle
    (p, p, p) has centroid ((p + p) + p) / 3 == p exactly in FP32."""
    return (rng.integers(-4096, 4096, (n, 3)) / 64.0).astype(np.float32)


@pytest.fixture(scope="module")
def learn(orc):
    if orc.learn_lbvh(np.zeros((2, 9), np.float32)) is None:
        pytest.skip("the reference walkthrough build (oracle/_ref) is not present")
    return orc.learn_lbvh


def test_lbvh_links_equal_reference_walkthrough_triangles(orc, scenes, learn):
    """Leaf order, child ids and depth of the CPU LBVH equal the reference's own
    LBVH walkthrough (learn/03_lbvh.cpp) on terrain, random and duplicate triangles."""
    rng = np.random.default_rng(77)
    a = rng.uniform(-50, 50, (20000, 3))
    rand = np.concatenate([a, a + rng.uniform(-5, 5, (20000, 3)), a + rng.uniform(-5, 5, (20000, 3))], 1)
    dup = np.repeat(rand[:300], 23, axis=0)
    for tris in (scenes.terrain(64, 3), rand.astype(np.float32), dup.astype(np.float32)):
        n = tris.shape[0]
        raw, ids, depth = orc.lbvh_nodes(0, tris)
        order, left, right, d = learn(tris)
        l2, r2 = _links_depth(raw, n)
        assert np.array_equal(ids, order)
        assert np.array_equal(l2, left) and np.array_equal(r2, right)
        assert depth == d


@pytest.mark.parametrize("kind", [1, 2])
def test_lbvh_links_equal_reference_walkthrough_points_gaussians(orc, learn, kind):
    """Points and Gaussian means take the triangle path's keys when each is posed as
    the triangle (p, p, p), whose centroid is p exactly for these coordinates."""
    rng = np.random.default_rng(80 + kind)
    pts = _dyadic_points(rng, 15000)
    pts[100:300] = pts[9]
    tris = np.concatenate([pts, pts, pts], 1).astype(np.float32)
    t = tris.reshape(-1, 3, 3)
    assert np.array_equal(((t[:, 0] + t[:, 1]) + t[:, 2]) / np.float32(3.0), pts)
    items = _items(kind, pts.astype(np.float64), rng)
    raw, ids, depth = orc.lbvh_nodes(kind, items)
    order, left, right, d = learn(tris)
    l2, r2 = _links_depth(raw, pts.shape[0])
    assert np.array_equal(ids, order)
    assert np.array_equal(l2, left) and np.array_equal(r2, right)
    assert depth == d
