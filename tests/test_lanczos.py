"""Block Lanczos for M^{1/2} Z (the Brownian-displacement driver over the
batched apply) and the tall-skinny block kernels it runs on.

CPU: the recurrence with numpy block operations on SPD matrices whose square
root numpy's eigh gives exactly (the error falls with tol; breakdown on a
rank-deficient block and a non-SPD operator are reported).
GPU: spme_block_gram / spme_block_update against numpy (FP32 and FP64,
uneven widths, ld > n, beta = 0 not reading Y); M^{1/2} Z of the engine's
full operator against the dense square root of the same operator, built
column by column through the batched apply (N = 100); the displacement's
deterministic part equals dt M F.
"""
import numpy as np
import pytest

from spme_amd import workload as W
from spme_amd.lanczos import BlockLanczos
from spme_amd.lib import SpmeError


class NumpyOps:
    """Block operations on host arrays: the recurrence's CPU check."""
    eps = 1e-7

    def __init__(self, n):
        self.n = n

    def gram(self, X, Y):
        return X @ Y.T

    def update(self, X, A, beta, Y):
        Y[...] = (0.0 if beta == 0 else beta * Y) + A.T @ X

    def empty(self, s):
        return np.empty((s, self.n))


def _spd(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    return (Q * lam) @ Q.T, (Q * np.sqrt(lam)) @ Q.T


@pytest.mark.parametrize("s,n", [(16, 400), (4, 250), (1, 120)])
def test_block_lanczos_matches_dense_sqrt(s, n):
    M, Mh = _spd(n, 0.2, 5.0, seed=s + n)
    Z = np.random.default_rng(7).standard_normal((s, n))
    ref = Z @ Mh.T
    errs = []
    for tol in (1e-3, 1e-6, 1e-10):
        bl = BlockLanczos(lambda X, Y: Y.__setitem__(Ellipsis, X @ M.T), NumpyOps(n), s,
                          tol=tol, max_iter=80)
        U, info = bl.sqrt(Z)
        assert info.converged
        assert 0.2 * 0.999 < info.min_eig and info.max_eig < 5.0 * 1.001
        errs.append(np.linalg.norm(U - ref) / np.linalg.norm(ref))
    # the stopping rule tracks the true error
    assert errs[0] < 1e-2 and errs[1] < 1e-5 and errs[2] < 1e-9, errs
    assert errs[0] > errs[1] > errs[2]


def test_block_lanczos_deflation_and_errors():
    n, s = 200, 4
    M, Mh = _spd(n, 0.5, 2.0, seed=3)
    op = (lambda X, Y: Y.__setitem__(Ellipsis, X @ M.T))
    # a rank-deficient block deflates: equal rows give equal results
    Z = np.random.default_rng(1).standard_normal((s, n))
    Z[3] = Z[1]
    U, info = BlockLanczos(op, NumpyOps(n), s, tol=1e-10).sqrt(Z)
    assert info.converged
    ref = Z @ Mh.T
    assert np.linalg.norm(U - ref) / np.linalg.norm(ref) < 1e-9
    # a Krylov space that fills the whole space ends invariant, and exact
    n2 = 40
    M2, Mh2 = _spd(n2, 0.1, 10.0, seed=4)
    Z2 = np.random.default_rng(2).standard_normal((16, n2))
    U2, info2 = BlockLanczos(lambda X, Y: Y.__setitem__(Ellipsis, X @ M2.T), NumpyOps(n2), 16,
                             tol=1e-14).sqrt(Z2)
    assert info2.invariant and info2.dimension == n2
    ref2 = Z2 @ Mh2.T
    assert np.linalg.norm(U2 - ref2) / np.linalg.norm(ref2) < 1e-11
    # an indefinite operator is reported
    Mneg = M - 1.5 * np.eye(n)
    Z = np.random.default_rng(2).standard_normal((s, n))
    bl = BlockLanczos(lambda X, Y: Y.__setitem__(Ellipsis, X @ Mneg.T), NumpyOps(n), s)
    with pytest.raises(SpmeError, match="positive definite"):
        bl.sqrt(Z)
    with pytest.raises(SpmeError, match="block width"):
        BlockLanczos(None, NumpyOps(n), 17)


# ---------------------------------------------------------------------------
def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_block_kernels_match_numpy(prec):
    torch = _torch()
    from spme_amd import engine as E
    from spme_amd.lanczos import HipBlockOps
    print(f"case: block gram/update {prec}", flush=True)
    P = E.SPME_FP32 if prec == "fp32" else E.SPME_FP64
    dt = torch.float32 if prec == "fp32" else torch.float64
    n, sx, sy, ld = 10007, 13, 16, 10040
    rng = np.random.default_rng(5)
    X = rng.standard_normal((sx, ld))
    Y = rng.standard_normal((sy, ld))
    A = rng.standard_normal((sx, sy))
    if prec == "fp32":
        X, Y = X.astype(np.float32).astype(np.float64), Y.astype(np.float32).astype(np.float64)
    dX = torch.tensor(X, device="cuda", dtype=dt)
    dY = torch.tensor(Y, device="cuda", dtype=dt)
    ops = HipBlockOps(P, n)
    G = ops.gram(dX[:, :n], dY[:, :n])
    Gref = X[:, :n] @ Y[:, :n].T
    assert np.abs(G - Gref).max() < 1e-11 * np.abs(Gref).max()
    tol = 1e-6 if prec == "fp32" else 1e-13
    ops.update(dX[:, :n], A, 1.5, dY[:, :n])
    ref = 1.5 * Y[:, :n] + A.T @ X[:, :n]
    got = dY.double().cpu().numpy()
    assert np.abs(got[:, :n] - ref).max() < tol * np.abs(ref).max()
    np.testing.assert_array_equal(got[:, n:], Y[:, n:])       # beyond n untouched
    dZ = torch.full((sy, ld), float("nan"), device="cuda", dtype=dt)
    ops.update(dX[:, :n], A, 0.0, dZ[:, :n])                   # beta = 0 reads nothing
    got = dZ.double().cpu().numpy()[:, :n]
    assert np.abs(got - A.T @ X[:, :n]).max() < tol * np.abs(A.T @ X[:, :n]).max()
    with pytest.raises(SpmeError, match="alias"):
        ops.update(dX[:, :n], A[:, :13], 0.0, dX[:, :n])


def _dense_operator(E, eng, N, dt):
    torch = _torch()
    n3 = 3 * N
    M = np.zeros((n3, n3))
    for c0 in range(0, n3, 16):
        c = min(16, n3 - c0)
        X = torch.zeros(c, n3, device="cuda", dtype=dt)
        X[torch.arange(c), torch.arange(c0, c0 + c)] = 1.0
        Y = torch.zeros_like(X)
        eng.apply(X, Y, parts=E.SPME_FULL)
        M[:, c0:c0 + c] = Y.double().cpu().numpy().T
    return M


@pytest.mark.gpu
def test_engine_sqrt_matches_dense_sqrt():
    torch = _torch()
    from spme_amd import engine as E
    from spme_amd.lanczos import engine_lanczos
    print("case: engine M^1/2 Z, N=100 phi=0.2 K=32 p=6 rc=4 fp64/fp32", flush=True)
    N = 100
    s = W.make_system(4242, N, 0.2)
    # xi = 0.6: the Nyquist planes, where the wave-space operator is not
    # exactly symmetric (index K/2 maps to +pi K/L), carry exp(-43)
    e64 = E.Engine(0.6, 32, 6, N, s.L, rc=4.0, precision=E.SPME_FP64, max_rhs=16)
    e64.set_positions(s.pos)
    M = _dense_operator(E, e64, N, torch.float64)
    assert np.abs(M - M.T).max() < 1e-11 * np.abs(M).max()    # symmetric
    lam, Q = np.linalg.eigh(M)
    assert lam[0] > 0.0                                       # positive definite
    Mh = (Q * np.sqrt(lam)) @ Q.T
    Z = W.forces(4243, N, 16)
    ref = Z @ Mh.T
    for eng, dt, tol, bar in [(e64, torch.float64, 1e-9, 1e-7)]:
        bl = engine_lanczos(eng, 16, tol=tol, max_iter=60)
        U, info = bl.sqrt(torch.tensor(Z, device="cuda", dtype=dt))
        err = np.linalg.norm(U.double().cpu().numpy() - ref) / np.linalg.norm(ref)
        print(f"  fp64: {info.iterations} iterations, change {info.change:.2e}, "
              f"error {err:.2e}", flush=True)
        assert info.converged and err < bar
        # Ritz values lie inside the spectrum
        assert lam[0] * (1 - 1e-9) <= info.min_eig and info.max_eig <= lam[-1] * (1 + 1e-9)
    e32 = E.Engine(0.6, 32, 6, N, s.L, rc=4.0, precision=E.SPME_FP32, max_rhs=16)
    e32.set_positions(s.pos)
    bl = engine_lanczos(e32, 16, tol=1e-5, max_iter=60)
    U, info = bl.sqrt(torch.tensor(Z, device="cuda", dtype=torch.float32))
    err = np.linalg.norm(U.double().cpu().numpy() - ref) / np.linalg.norm(ref)
    print(f"  fp32: {info.iterations} iterations, error {err:.2e}", flush=True)
    assert info.converged and err < 1e-4
    e64.close()
    e32.close()


@pytest.mark.gpu
def test_brownian_displacement_parts():
    torch = _torch()
    from spme_amd import engine as E
    from spme_amd.lanczos import brownian_displacements, engine_lanczos
    print("case: brownian displacements, N=2000 phi=0.2 K=48 p=6 rc=6 xi=0.5 fp64",
          flush=True)
    N = 2000
    s = W.make_system(99, N, 0.2)
    # erfc(xi rc) = 2e-5 and exp(-19) at the Nyquist planes: an accurate, SPD
    # operator (a cutoff of erfc ~ 1e-2 leaves negative eigenvalues)
    eng = E.Engine(0.5, 48, 6, N, s.L, rc=6.0, precision=E.SPME_FP64, max_rhs=16)
    eng.set_positions(s.pos)
    F = torch.tensor(W.forces(100, N, 16), device="cuda")
    Z = torch.tensor(W.forces(101, N, 16), device="cuda")
    dt, kT = 1e-3, 1.0
    dX, info = brownian_displacements(eng, F, Z, dt, kT, tol=1e-8)
    assert info.converged
    MF = eng.apply(F, parts=E.SPME_FULL)
    U, _ = engine_lanczos(eng, 16, tol=1e-8).sqrt(Z)
    ref = dt * MF + np.sqrt(2 * kT * dt) * U
    err = float((dX - ref).abs().max() / ref.abs().max())
    assert err < 1e-10, err
    # M^{1/2} Z has the norms M prescribes: |M^{1/2} z|^2 = z^T M z
    MZ = eng.apply(Z, parts=E.SPME_FULL)
    q1 = (U * U).sum(dim=1)
    q2 = (Z * MZ).sum(dim=1)
    assert float(((q1 - q2).abs() / q2).max()) < 1e-6
    eng.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_tall_block_kernels_match_numpy(prec):
    """The tall form (X up to 4096 rows: every earlier block of the basis
    at once) against numpy in FP64, deterministic, with its limits."""
    torch = _torch()
    from spme_amd import engine as E
    from spme_amd.lanczos import HipBlockOps
    print(f"case: tall block gram/update {prec}", flush=True)
    P = E.SPME_FP32 if prec == "fp32" else E.SPME_FP64
    dt = torch.float32 if prec == "fp32" else torch.float64
    n, ld = 10007, 10040
    ops = HipBlockOps(P, n)
    rng = np.random.default_rng(9)
    for sx, sy in [(17, 16), (100, 7), (1000, 16)]:
        X = rng.standard_normal((sx, ld))
        Y = rng.standard_normal((sy, ld))
        A = rng.standard_normal((sx, sy))
        if prec == "fp32":
            X = X.astype(np.float32).astype(np.float64)
            Y = Y.astype(np.float32).astype(np.float64)
        dX = torch.tensor(X, device="cuda", dtype=dt)
        dY = torch.tensor(Y, device="cuda", dtype=dt)
        G1 = ops.gram_dev(dX[:, :n], dY[:, :n]).cpu().numpy()
        G2 = ops.gram_dev(dX[:, :n], dY[:, :n]).cpu().numpy()
        np.testing.assert_array_equal(G1, G2)                     # deterministic
        Gref = X[:, :n] @ Y[:, :n].T
        assert np.abs(G1 - Gref).max() < 1e-11 * np.abs(Gref).max(), (sx, sy)
        tol = 1e-6 if prec == "fp32" else 1e-13
        ops.update(dX[:, :n], A, 1.5, dY[:, :n])
        ref = 1.5 * Y[:, :n] + A.T @ X[:, :n]
        got = dY.double().cpu().numpy()
        assert np.abs(got[:, :n] - ref).max() < tol * np.abs(ref).max(), (sx, sy)
        np.testing.assert_array_equal(got[:, n:], Y[:, n:])
    # an update whose output rows lie inside X's span is refused
    with pytest.raises(SpmeError, match="alias"):
        ops.update(dX[:40, :n], A[:40, :4], 0.0, dX[30:34, :n])
    big = torch.zeros(4097, 8, device="cuda", dtype=dt)
    small = torch.zeros(2, 8, device="cuda", dtype=dt)
    with pytest.raises(SpmeError, match="block widths"):
        HipBlockOps(P, 8).gram_dev(big, small)
