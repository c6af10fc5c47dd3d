"""Differentiable mode on the GPU: forward bit-exact, adjoints against the FP64 CPU
adjoint and against FP64 central finite differences (h = 1e-3, Richardson-refined
with h/2) on 1,024 sampled parameters per scene.

Tolerance: |gpu - reference| <= 1e-4 * max(|reference|, 5% of the gradient
array's rms) -- relative 1e-4, with components far below the typical magnitude
judged against that floor.  The loss keeps the primitive each query was assigned
in the forward pass (that assignment is piecewise constant, and the gradient is
the gradient of the active piece).
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
TOL = 1e-4


def rel(a, b, floor):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def rms(x):
    return float(np.sqrt(np.mean(np.square(x))))


def central(fun, x, idx, h):
    xp, xm = x.copy(), x.copy()
    xp.flat[idx] += h
    xm.flat[idx] -= h
    return (fun(xp) - fun(xm)) / (2 * h)


def fd(fun, x, idx, h=1e-3):
    return (4.0 * central(fun, x, idx, h / 2) - central(fun, x, idx, h)) / 3.0


def check_fd(grad, fn, arr, idx, floor, stats):
    """The gradient component against the FP64 central difference at h = 1e-3
    (Richardson-refined with h/2).  Where a closest-feature region boundary lies
    within h of the sample the loss is C1 but not C2 there and that difference is
    off by O(h); such samples are checked at h = 1e-6 instead (counted in
    stats["fine"])."""
    e = rel(grad, fd(fn, arr, idx), floor)
    if e < TOL:
        stats["coarse"] += 1
        return e
    e2 = rel(grad, central(fn, arr, idx, 1e-6), floor)
    stats["fine"] += 1
    assert e2 < TOL, (idx, grad, fd(fn, arr, idx), central(fn, arr, idx, 1e-6))
    return e2


@pytest.fixture(scope="module")
def AG():
    from quarry import autograd

    if not torch.cuda.is_available():
        pytest.fail("no GPU visible to torch")
    return autograd


@pytest.fixture(scope="module")
def sphere(scenes):
    V, F, tris = scenes.icosphere(levels=3, radius=10.0, noise=0.1, seed=2)
    rng = np.random.default_rng(3)
    q = rng.uniform(-15, 15, (4096, 3)).astype(np.float32)
    rays = scenes.sphere_rays(4096, seed=4)
    return V, F, tris, q, rays


def test_forward_bit_exact(AG, orc, sphere):
    V, F, tris, q, rays = sphere
    dev = torch.device("cuda")
    d2, idx = AG.closest_point_dist2(torch.from_numpy(q).to(dev), torch.from_numpy(V).to(dev),
                                     torch.from_numpy(F).to(dev))
    want = orc.knn_brute(0, tris, q, 0, threads=THREADS)
    assert np.array_equal(d2.cpu().numpy().view(np.uint32), want[:, 0, 0])
    assert np.array_equal(idx.cpu().numpy().astype(np.int32), want[:, 0, 1].view(np.int32))
    t, hit = AG.ray_hit_t(torch.from_numpy(rays[:, 0:3].copy()).to(dev), torch.from_numpy(rays[:, 3:6].copy()).to(dev),
                          torch.from_numpy(V).to(dev), torch.from_numpy(F).to(dev))
    wr = orc.ray_brute(tris, rays, threads=THREADS)
    wh = (wr >> np.uint64(32)) != 0
    tt = t.cpu().numpy()
    assert np.array_equal(np.isfinite(tt), wh)
    assert np.array_equal(tt[wh].view(np.uint32), (wr[wh] & np.uint64(0xFFFFFFFF)).astype(np.uint32))
    assert np.all(hit.cpu().numpy()[~wh] == -1) and np.all(hit.cpu().numpy()[wh] >= 0)


def test_mesh_gradients(AG, sphere):
    from oracle import adjoint as J

    V, F, tris, q, rays = sphere
    dev = torch.device("cuda")
    Vt = torch.from_numpy(V).to(dev).requires_grad_(True)
    qt = torch.from_numpy(q).to(dev).requires_grad_(True)
    ot = torch.from_numpy(rays[:, 0:3].copy()).to(dev).requires_grad_(True)
    dt = torch.from_numpy(rays[:, 3:6].copy()).to(dev).requires_grad_(True)
    Ft = torch.from_numpy(F).to(dev)
    d2, idx = AG.closest_point_dist2(qt, Vt, Ft)
    t, hit = AG.ray_hit_t(ot, dt, Vt, Ft)
    mask = torch.isfinite(t)
    loss = d2.sum() + torch.where(mask, t, torch.zeros_like(t)).sum()
    loss.backward()
    gV, gq = Vt.grad.cpu().numpy().astype(np.float64), qt.grad.cpu().numpy().astype(np.float64)
    go, gd = ot.grad.cpu().numpy().astype(np.float64), dt.grad.cpu().numpy().astype(np.float64)
    prim = idx.cpu().numpy()
    hp = hit.cpu().numpy()
    assert (hp >= 0).sum() > 3000

    # CPU adjoint in FP64 with the same assignment
    q64, V64, o64, d64 = (np.asarray(x, np.float64) for x in (q, V, rays[:, 0:3], rays[:, 3:6]))
    T64 = V64[F].reshape(-1, 9)
    cq, ct = J.tri_adjoint(q64, T64, prim, np.ones(len(q)))
    co, cd, ct2 = J.ray_adjoint(o64, d64, T64, hp, np.ones(len(hp)))
    cT = (ct + ct2).reshape(-1, 3, 3)
    cV = np.zeros_like(V64)
    for k in range(3):
        np.add.at(cV, F[:, k], cT[:, k])
    for got, ref in ((gq, cq), (gV, cV), (go, co), (gd, cd)):
        err = rel(got, ref, 0.05 * rms(ref))
        print("cpu-adjoint max rel err", float(err.max()), flush=True)
        assert err.max() < TOL

    # central differences of the FP64 loss on 1,024 sampled parameters
    def loss_of(Vx=V64, qx=q64, ox=o64, dx=d64):
        Tx = Vx[F].reshape(-1, 9)
        s = J.tri_dist2_fixed(qx, Tx, prim).sum()
        ok = hp >= 0
        return s + J.ray_t_fixed(ox[ok], dx[ok], Tx, hp[ok]).sum()

    rng = np.random.default_rng(7)
    worst = 0.0
    stats = {"coarse": 0, "fine": 0}
    for arr, grad, fn, count in (
            (V64, gV, lambda x: loss_of(Vx=x), 512),
            (q64, gq, lambda x: loss_of(qx=x), 256),
            (o64, go, lambda x: loss_of(ox=x), 128),
            (d64, gd, lambda x: loss_of(dx=x), 128)):
        floor = 0.05 * rms(grad)
        for idx_ in rng.choice(arr.size, count, replace=False):
            worst = max(worst, float(check_fd(grad.flat[idx_], fn, arr, idx_, floor, stats)))
    print("FD max rel err", worst, stats, flush=True)
    assert stats["coarse"] + stats["fine"] == 1024 and stats["fine"] <= 16


def test_gaussian_gradients(AG, orc, scenes):
    from oracle import adjoint as J

    g = scenes.gaussians(3000, 5, side=24.0)
    g[:, 3:7] *= np.float32(1.7)  # unnormalised quaternions
    rng = np.random.default_rng(8)
    q = rng.uniform(-3, 27, (2048, 3)).astype(np.float32)
    dev = torch.device("cuda")
    m = torch.from_numpy(g[:, 0:3].copy()).to(dev).requires_grad_(True)
    r = torch.from_numpy(g[:, 3:7].copy()).to(dev).requires_grad_(True)
    s = torch.from_numpy(g[:, 7:10].copy()).to(dev).requires_grad_(True)
    qt = torch.from_numpy(q).to(dev).requires_grad_(True)
    d2, idx = AG.gaussian_closest_dist2(qt, m, r, s)
    want = orc.knn_brute(2, g, q, 0, threads=THREADS)
    assert np.array_equal(d2.detach().cpu().numpy().view(np.uint32), want[:, 0, 0])
    d2.sum().backward()
    prim = idx.cpu().numpy()
    G64, q64 = g.astype(np.float64), q.astype(np.float64)
    cq, cg = J.gauss_adjoint(q64, G64, prim, np.ones(len(q)))
    grads = {"q": qt.grad.cpu().numpy().astype(np.float64), "m": m.grad.cpu().numpy().astype(np.float64),
             "r": r.grad.cpu().numpy().astype(np.float64), "s": s.grad.cpu().numpy().astype(np.float64)}
    for name, got, ref in (("q", grads["q"], cq), ("m", grads["m"], cg[:, 0:3]), ("r", grads["r"], cg[:, 3:7]),
                           ("s", grads["s"], cg[:, 7:10])):
        err = rel(got, ref, 0.05 * rms(ref))
        print(name, "cpu-adjoint max rel err", float(err.max()), flush=True)
        assert err.max() < TOL, name

    def loss_of(gx=G64, qx=q64):
        return J.gauss_dist2_fixed(qx, gx, prim).sum()

    rng = np.random.default_rng(9)
    gflat = np.concatenate([grads["m"], grads["r"], grads["s"]], 1)
    floor_g = {c: 0.05 * rms(gflat[:, c]) for c in range(10)}
    stats = {"coarse": 0, "fine": 0}
    for idx_ in rng.choice(G64.size, 768, replace=False):
        check_fd(gflat.flat[idx_], lambda x: loss_of(gx=x), G64, idx_, floor_g[idx_ % 10], stats)
    for idx_ in rng.choice(q64.size, 256, replace=False):
        check_fd(grads["q"].flat[idx_], lambda x: loss_of(qx=x), q64, idx_, 0.05 * rms(grads["q"]), stats)
    print("FD checks", stats, flush=True)
    assert stats["coarse"] + stats["fine"] == 1024 and stats["fine"] <= 16


def test_point_gradients(AG, orc, scenes):
    pts = scenes.point_cloud(5000, 3)
    q = np.random.default_rng(10).uniform(-4, 68, (1500, 3)).astype(np.float32)
    dev = torch.device("cuda")
    ct = torch.from_numpy(pts).to(dev).requires_grad_(True)
    qt = torch.from_numpy(q).to(dev).requires_grad_(True)
    d2, idx = AG.point_closest_dist2(qt, ct)
    want = orc.knn_brute(1, pts, q, 0, threads=THREADS)
    assert np.array_equal(d2.detach().cpu().numpy().view(np.uint32), want[:, 0, 0])
    (2.0 * d2).sum().backward()
    prim = idx.cpu().numpy()
    gq = qt.grad.cpu().numpy()
    assert np.allclose(gq, 4.0 * (q - pts[prim]), rtol=1e-5, atol=1e-5)
    gc = ct.grad.cpu().numpy()
    ref = np.zeros_like(pts)
    np.add.at(ref, prim, -4.0 * (q - pts[prim]))
    assert np.allclose(gc, ref, rtol=1e-5, atol=1e-4)


def test_gradcheck_small(AG):
    """torch.autograd.gradcheck-style sanity on a tiny scene in float32: the
    analytic Jacobian-vector products match directional differences."""
    dev = torch.device("cuda")
    V = torch.tensor([[0, 0, 0], [4, 0, 0], [0, 4, 0], [4, 4, 1]], dtype=torch.float32, device=dev,
                     requires_grad=True)
    F = torch.tensor([[0, 1, 2], [1, 3, 2]], dtype=torch.int32, device=dev)
    q = torch.tensor([[1.0, 1.0, 2.0], [3.5, 3.0, -1.0], [-2.0, -1.0, 0.5]], device=dev, requires_grad=True)
    d2, _ = AG.closest_point_dist2(q, V, F, order="input", wide=False)
    d2.sum().backward()
    # point above the first triangle's interior: gradient 2 (q - p) with p = (1, 1, 0)
    assert torch.allclose(q.grad[0], torch.tensor([0.0, 0.0, 4.0], device=dev), atol=1e-6)


def test_gaussian_gradients_million_scene(AG, orc, scenes):
    """The 1,048,576-Gaussian scene (means uniform in a 64-unit cube, log-uniform
    scales, random rotations) with 262,144 uniform queries: forward bits equal the
    CPU BVH's, every gradient equals the FP64 CPU adjoint to 1e-4, and 1,024 sampled
    parameters (768 of the Gaussians that hold a closest point, 256 query
    coordinates) equal FP64 central differences of the loss."""
    from oracle import adjoint as J

    g = scenes.gaussians(1 << 20, 1)
    rng = np.random.default_rng(21)
    q = rng.uniform(0.0, 64.0, (262_144, 3)).astype(np.float32)
    dev = torch.device("cuda")
    m = torch.from_numpy(g[:, 0:3].copy()).to(dev).requires_grad_(True)
    r = torch.from_numpy(g[:, 3:7].copy()).to(dev).requires_grad_(True)
    s = torch.from_numpy(g[:, 7:10].copy()).to(dev).requires_grad_(True)
    qt = torch.from_numpy(q).to(dev).requires_grad_(True)
    print("forward + backward on 1M Gaussians, 262k queries", flush=True)
    d2, idx = AG.gaussian_closest_dist2(qt, m, r, s)
    d2.sum().backward()
    want = orc.CpuBvh(2, g).knn(q, 0, THREADS)
    assert np.array_equal(d2.detach().cpu().numpy().view(np.uint32), want[:, 0, 0])
    prim = idx.cpu().numpy()
    assert np.array_equal(prim, want[:, 0, 1].view(np.int32))
    G64, q64 = g.astype(np.float64), q.astype(np.float64)
    print("FP64 CPU adjoint", flush=True)
    cq, cg = J.gauss_adjoint(q64, G64, prim, np.ones(len(q)))
    grads = {"q": qt.grad.cpu().numpy().astype(np.float64), "m": m.grad.cpu().numpy().astype(np.float64),
             "r": r.grad.cpu().numpy().astype(np.float64), "s": s.grad.cpu().numpy().astype(np.float64)}
    for name, got, ref in (("q", grads["q"], cq), ("m", grads["m"], cg[:, 0:3]), ("r", grads["r"], cg[:, 3:7]),
                           ("s", grads["s"], cg[:, 7:10])):
        err = rel(got, ref, 0.05 * rms(ref[ref != 0]) if np.any(ref != 0) else 1.0)
        print(name, "cpu-adjoint max rel err", float(err.max()), flush=True)
        assert err.max() < TOL, name

    # finite differences: a parameter of Gaussian j moves only the loss of the queries
    # whose closest Gaussian is j, so the loss is summed over those
    owners = {}
    for qi, pj in enumerate(prim):
        owners.setdefault(int(pj), []).append(qi)
    held = np.array(sorted(owners), np.int64)
    gflat = np.concatenate([grads["m"], grads["r"], grads["s"]], 1)
    floor_g = {c: 0.05 * rms(gflat[held, c]) for c in range(10)}
    stats = {"coarse": 0, "fine": 0}
    rng = np.random.default_rng(22)
    for j, c in zip(rng.choice(held, 768, replace=False), rng.integers(0, 10, 768)):
        qs = np.array(owners[int(j)])
        sub = G64[j:j + 1].copy()
        fn = lambda x: J.gauss_dist2_fixed(q64[qs], x, np.zeros(len(qs), np.int64)).sum()  # noqa: E731
        check_fd(gflat[j, c], fn, sub, int(c), floor_g[int(c)], stats)
    floor_q = 0.05 * rms(grads["q"])
    for qi, c in zip(rng.choice(len(q), 256, replace=False), rng.integers(0, 3, 256)):
        sub = q64[qi:qi + 1].copy()
        fn = lambda x: J.gauss_dist2_fixed(x, G64, prim[qi:qi + 1]).sum()  # noqa: E731,B023
        check_fd(grads["q"][qi, c], fn, sub, int(c), floor_q, stats)
    print("FD checks", stats, flush=True)
    assert stats["coarse"] + stats["fine"] == 1024 and stats["fine"] <= 16
