"""Parameter selection: the analytic estimate, the learned network's CPU
restatement (pinned by finite differences), the GPU forward/backward
kernels against that restatement, and training."""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as MO
from spme_amd import params as PR


def test_finite_difference_gradient_of_restatement():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.3, MO.NPARAM)
    x = rng.normal(size=(7, 4))
    t = rng.normal(size=(7, 4))
    mask = np.array([1.0, 0.5, 2.0, 1.0])
    g = MO.grad(w, x, t, mask)
    idx = rng.choice(MO.NPARAM, 60, replace=False)
    for i in idx:
        e = np.zeros(MO.NPARAM)
        e[i] = 1e-6
        fd = (MO.loss(w + e, x, t, mask) - MO.loss(w - e, x, t, mask)) / 2e-6
        assert abs(fd - g[i]) < 1e-7 + 1e-6 * abs(fd)


def test_layout_matches_kernel_parameter_count():
    assert MO.NPARAM == 4 * 64 + 64 + 64 * 64 + 64 + 64 * 4 + 4


def test_real_error_estimate_matches_measurement():
    # measured truncation of the real-space sum (restated reference, h = 0.84
    # box of 3125 spheres at phi = 0.2, rc = 4) divided by rms(u) = 2.12
    for xi, meas in [(0.5, 3.84e-2), (0.7, 5.67e-3), (1.0, 1.56e-5)]:
        est = PR.real_error_abs(xi, 4.0, 0.2) / 2.12
        assert 0.8 < est / meas < 1.25, (xi, est, meas)


def test_analytic_selection_meets_estimate_and_constraints():
    for n, phi, tol in [(1000, 0.1, 1e-4), (100000, 0.2, 1e-5), (500000, 0.3, 1e-6)]:
        a = PR.analytic_select(n, phi, tol)
        L = PR.box_length(n, phi)
        assert a.K in PR.SMOOTH_K and a.p in PR.ORDERS and 4 * a.p < a.K
        assert a.rc <= L / 3.0 + 1e-9
        assert PR.estimate_error(a.xi, a.K, a.p, a.rc, phi, L) <= tol * 1.01
    # fixed mesh and order: only xi and rc move
    a = PR.analytic_select(200000, 0.2, 1e-4, fixed_K=192, fixed_p=8)
    assert (a.K, a.p) == (192, 8) and 0.4 < a.xi < 0.6 and 6 < a.rc < 9


def test_labels_pick_fastest_feasible():
    from spme_amd import mlp
    runs = {"sys": np.array([0, 0, 0]), "n": np.array([1000] * 3),
            "phi": np.array([0.1] * 3), "xi": np.array([0.3, 0.4, 0.5]),
            "K": np.array([32, 48, 64]), "p": np.array([6, 6, 8]),
            "rc": np.array([8.0, 7.0, 6.0]), "err": np.array([1e-5, 2e-4, 5e-6]),
            "time": np.array([2.0, 1.0, 3.0])}
    X, T, info = mlp.labels_from_runs(runs, tols=(1e-3, 1e-5, 1e-6))
    assert [i[2] for i in info] == [1, 0]          # tol 1e-6 infeasible
    L = PR.box_length(1000, 0.1)
    assert np.allclose(T[0], mlp.targets(0.4, 48, 6, 7.0, L))
    assert np.isclose(np.exp(T[0][1]), L / 48)     # the mesh spacing, not K
    # the tolerance enters in absolute terms: tol * sqrt(1 + 0.66 phi L)
    assert np.isclose(10 ** X[0][2], 1e-3 * np.sqrt(1 + 0.66 * 0.1 * L))


def test_mesh_labels_interpolate_the_crossing():
    """The smallest cutoff at which the error stays within the target,
    log-interpolated between grid cutoffs; the xi reaching it first wins."""
    from spme_amd import mlp
    rc = np.array([2.0, 4.0, 6.0, 8.0])
    e1 = np.array([1e-1, 1e-3, 1e-5, 1e-5])   # xi = 0.5: crosses 1e-4 between 4 and 6
    e2 = np.array([1e-2, 1e-4 * 1.5, 2e-4, 2e-4])  # xi = 0.9: floor above the target
    runs = {k: np.concatenate([v, v]) for k, v in
            dict(sys=np.zeros(4, int), K=np.full(4, 64), p=np.full(4, 6),
                 L=np.full(4, 50.0), phi=np.full(4, 0.2), rc=rc).items()}
    runs["xi"] = np.array([0.5] * 4 + [0.9] * 4)
    runs["err"] = np.concatenate([e1, e2])
    X, T, info = mlp.mesh_labels(runs, tols=(1.5e-4,), margin=1.5)
    assert len(X) == 1
    assert np.isclose(np.exp(T[0][0]), 0.5)
    assert np.isclose(np.exp(T[0][1]), 5.0)          # halfway in log error from 1e-3 to 1e-5
    assert np.allclose(X[0], mlp.mesh_features(0.2, 1.5e-4, 50.0 / 64, 6, 50.0))
    assert mlp.mesh_for(191.1, 0.6) == 256 and mlp.mesh_for(100.0, 0.8) == 128
    assert mlp.mesh_for(100.0, 0.95) == 96 and mlp.mesh_for(100.0, 0.5, k_max=128) == 128


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_restatement():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from spme_amd import mlp
    rng = np.random.default_rng(1)
    w = rng.normal(0, 0.3, MO.NPARAM).astype(np.float32)
    x = rng.normal(size=(301, 4)).astype(np.float32)
    t = rng.normal(size=(301, 4)).astype(np.float32)
    net = mlp.Net(w, np.zeros(4), np.ones(4), np.zeros(4), np.ones(4))
    y = net.forward_raw(x)
    yr, _, _ = MO.forward(w.astype(np.float64), x.astype(np.float64))
    assert np.abs(y - yr).max() < 1e-5 * max(1.0, np.abs(yr).max())
    # one training step with lr = 0 leaves w; gradient via the backward kernel
    from spme_amd.lib import load
    import ctypes
    L = load()
    dev = lambda a: torch.tensor(a, device="cuda").contiguous()
    xd, td, md = dev(x), dev(t), dev(np.array([1, 0.5, 2, 1], np.float32))
    yd = torch.empty((301, 4), device="cuda")
    h1 = torch.empty((301, 64), device="cuda")
    h2 = torch.empty_like(h1)
    work = torch.empty(301 * (4 + 128), device="cuda")
    g = torch.empty(MO.NPARAM, device="cuda")
    p = lambda a: ctypes.c_void_p(a.data_ptr())
    assert L.spme_mlp_forward(p(net.w), p(xd), 301, p(yd), p(h1), p(h2), None) == 0
    assert L.spme_mlp_backward(p(net.w), p(xd), p(yd), p(td), p(md), 301, p(h1), p(h2),
                               p(work), p(g), None) == 0
    torch.cuda.synchronize()
    gr = MO.grad(w.astype(np.float64), x.astype(np.float64), t.astype(np.float64),
                 np.array([1, 0.5, 2, 1.0]))
    gg = g.cpu().numpy()
    assert np.abs(gg - gr).max() < 1e-4 * np.abs(gr).max()


@pytest.mark.gpu
def test_gpu_training_fits_a_smooth_map():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from spme_amd import mlp
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, size=(512, 4))
    T = np.stack([np.sin(X[:, 0]) + X[:, 1], X[:, 2] * X[:, 3], np.cos(X[:, 1]),
                  X.sum(1) / 4], 1)
    net = mlp.Net(mlp.init_params(0), np.zeros(4), np.ones(4), np.zeros(4), np.ones(4))
    hist = net.train(X, T, epochs=1500, lr=3e-3, log_every=500)
    assert hist[-1][1] < 0.02 * hist[0][1] or hist[-1][1] < 1e-3
    # bitwise reproducible
    net2 = mlp.Net(mlp.init_params(0), np.zeros(4), np.ones(4), np.zeros(4), np.ones(4))
    net2.train(X, T, epochs=1500, lr=3e-3)
    assert torch.equal(net.w, net2.w)
