"""Topology updates (delete / insert in the middle / append / move).

The new scene is composed on the device from the old tree's items plus the
uploaded new and moved primitives, then rebuilt: the resulting tree must equal a
fresh build of the new scene byte for byte, and queries must equal brute force.
Edits follow the reference's contract (src/gpu/gpu_bvh.h:182-213).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def A(qabi):
    if qabi.describe() is None:
        pytest.fail("no HIP device visible to the GPU tests")
    return qabi


def _edit(rng, old, make, n_del, n_ins, n_app, n_move):
    n = old.shape[0]
    deleted = np.sort(rng.choice(n, n_del, replace=False)).astype(np.int32)
    surv = np.delete(old, deleted, axis=0)
    new_items = make(n_ins + n_app)
    tail_beg = surv.shape[0] + n_ins
    inserted = np.sort(rng.choice(tail_beg, n_ins, replace=False)).astype(np.int32)
    out = np.empty((tail_beg + n_app, old.shape[1]), np.float32)
    mask = np.zeros(tail_beg, bool)
    mask[inserted] = True
    out[:tail_beg][mask] = new_items[:n_ins]
    out[:tail_beg][~mask] = surv
    out[tail_beg:] = new_items[n_ins:]
    cand = np.nonzero(~mask)[0]
    moved = np.sort(rng.choice(cand, min(n_move, cand.size), replace=False)).astype(np.int32)
    out[moved] += np.float32(0.37)
    return out, deleted, inserted, moved, n_app


@pytest.mark.parametrize("kind,builder", [(0, 0), (0, 1), (1, 0), (2, 0)])
def test_topology_rounds(A, orc, scenes, kind, builder):
    rng = np.random.default_rng(10 + kind + 5 * builder)
    if kind == 0:
        def make(m):
            a = rng.uniform(-20, 20, (m, 3))
            return np.concatenate([a, a + rng.normal(0, 1, (m, 3)), a + rng.normal(0, 1, (m, 3))], 1).astype(np.float32)
        items = make(6000)
    elif kind == 1:
        def make(m):
            return rng.uniform(-20, 20, (m, 3)).astype(np.float32)
        items = make(8000)
    else:
        def make(m):
            g = scenes.gaussians(m, int(rng.integers(1 << 30)), side=40.0)
            g[:, 0:3] -= np.float32(20.0)
            return g
        items = make(5000)
    b = A.Bvh(kind, items, builder)
    b.build_wide()
    q = rng.uniform(-24, 24, (600, 3)).astype(np.float32)
    for rnd, (nd, ni, na, nm) in enumerate(((60, 0, 0, 0), (0, 40, 0, 0), (0, 0, 75, 0), (0, 0, 0, 90),
                                             (120, 50, 30, 40), (3000, 10, 5, 5))):
        print(f"kind={kind} builder={builder} round={rnd}", flush=True)
        new, deleted, inserted, moved, appended = _edit(rng, items, make, nd, ni, na, nm)
        t = b.update(new, deleted, moved, inserted, appended)
        assert t.deleted == nd and t.inserted == ni and t.appended == na
        assert b.validate(new) == ""
        fresh = A.Bvh(kind, new, builder)
        if nd + ni + na > 0:
            assert b.tree_hash() == fresh.tree_hash(), rnd
        want = orc.knn_brute(kind, new, q, 6, threads=THREADS)
        for wide in (False, True):
            assert np.array_equal(b.knn(q, 6, wide=wide)[0], want), (rnd, wide)
        items = new


def test_topology_edit_errors(A):
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1, 1, (100, 3)).astype(np.float32)
    b = A.Bvh(1, pts)
    with pytest.raises(ValueError):
        b.update(pts, deleted=np.array([5, 5], np.int32))        # not unique
    with pytest.raises(ValueError):
        b.update(pts[:98], deleted=np.array([1, 200], np.int32))  # out of range
    with pytest.raises(ValueError):
        b.update(pts, inserted=np.array([3], np.int32))           # count mismatch
    with pytest.raises(ValueError):
        b.update(pts, moved=np.array([100], np.int32))            # moved out of range
    assert b.validate(pts) == ""
