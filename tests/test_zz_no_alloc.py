"""qr_knn_device_noalloc (kNN with its scratch in a fixed device array) against
qr_knn_device and the CPU oracle, through the C ABI: two calls in a row on one
tree, the second with a smaller query set, in input and Morton order, on binary
and 4-wide nodes; then the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = 8


def test_noalloc_knn_matches_allocating_entry_point(qabi, orc, scenes):
    A = qabi
    if A.describe() is None:
        pytest.fail("no HIP device visible to the GPU tests")
    L = A.lib()
    tris = scenes.terrain(96, 4)
    b = A.Bvh(0, tris)
    b.build_wide()
    rng = np.random.default_rng(5)
    ms = ctypes.c_double()
    for n, k in ((70000, 8), (3000, 8), (1200, 1), (40000, 16)):
        q = rng.uniform(-6, 102, (n, 3)).astype(np.float32)
        q[:, 2] = rng.uniform(-12, 12, n).astype(np.float32)
        dq = A.DeviceArray(q)
        want = orc.knn_brute(0, tris, q, k, threads=THREADS)
        for wide in (0, 1):
            for order in (A.ORDER_INPUT, A.ORDER_MORTON):
                print(f"noalloc n={n} k={k} wide={wide} order={order}", flush=True)
                out_a = A.DeviceArray.empty((n, k, 2), np.uint32)
                out_b = A.DeviceArray.empty((n, k, 2), np.uint32)
                A.check(L.qr_knn_device(b.h, dq.ptr, n, k, out_a.ptr, order, wide, ctypes.byref(ms)))
                A.check(L.qr_knn_device_noalloc(b.h, dq.ptr, n, k, out_b.ptr, order, wide))
                got_a, got_b = out_a.download(), out_b.download()
                assert np.array_equal(got_b, got_a), (n, k, wide, order)
                assert np.array_equal(got_b, want), (n, k, wide, order)


def test_noalloc_knn_rejects_sizes_beyond_its_scratch(qabi):
    A = qabi
    L = A.lib()
    b = A.Bvh(1, np.random.default_rng(1).uniform(-1, 1, (100, 3)).astype(np.float32))
    limit = 1 << 20
    # the size is checked before any pointer is read, so no buffers are needed
    with pytest.raises(A.QuarryError):
        A.check(L.qr_knn_device_noalloc(b.h, None, limit + 1, 4, None, A.ORDER_MORTON, 0))
    with pytest.raises(A.QuarryError):
        A.check(L.qr_knn_device_noalloc(b.h, None, 10, 0, None, A.ORDER_MORTON, 0))
    # an empty query set is fine
    A.check(L.qr_knn_device_noalloc(b.h, None, 0, 4, None, A.ORDER_MORTON, 0))
