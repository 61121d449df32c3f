"""GPU parity of the HIP engine against the CPU oracle (restated reference
SPME, old-version/spme.c + spme_kernel.c) through the C ABI.

FP64 bar: relative max-norm difference <= 1e-12 (same algorithm, different
summation order and fused multiply-adds).  FP32 bar: <= 2e-5 relative
(FP32 mesh, spreading, FFT and interpolation; FP64 tables).
"""
import sys

import numpy as np
import pytest

from spme_amd import workload as W

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-5


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


@pytest.fixture(scope="module")
def E():
    from spme_amd import engine
    return engine


def _system(seed, n, phi):
    s = W.make_system(seed, n, phi)
    return s.pos, s.L


def test_dropin_matches_oracle_c1(oracle, E, torch):
    """Config 1 through the reference's own surface, with the reference
    driver's alpha=2.1, beta=2.2 (old-version/main.c:224-245)."""
    _say("case: drop-in compute_spme, N=1000 phi=0.1 K=32 p=6")
    pos, L = _system(1001, 1000, 0.1)
    xi = E.cutoff2xi(5.0, 1e-6, L)
    f = W.forces(1001, 1000, 2)
    v0 = W.forces(1002, 1000, 2)
    ref = oracle.compute_spme(xi, 32, 6, L, pos, f, alpha=2.1, beta=2.2, u=v0)
    d = E.DropIn(xi, 32, 6, 1000, L)
    v = v0.copy()
    d.compute(pos, 2, 2.1, f, 3000, 2.2, v, 3000)
    assert _rel(v, ref) < FP64_TOL
    # pos == NULL reuses the P data
    v2 = v0.copy()
    d.compute(None, 2, 2.1, f, 3000, 2.2, v2, 3000)
    np.testing.assert_array_equal(v2, v)
    d.close()


def test_dropin_batches_many_rhs(oracle, E, torch):
    """compute_spme with 20 right-hand sides, padded ldin/ldout and alpha,
    beta: the drop-in runs them as batched applies (16 + 4 here); the result
    equals the oracle and, bitwise, one compute_spme call per right-hand
    side (the reference's loop, old-version/spme.c:631-702)."""
    _say("case: drop-in compute_spme, 20 RHS in batches, padded strides")
    N, K, p, R, ld = 600, 32, 6, 20, 1808
    pos, L = _system(61, N, 0.1)
    xi = 0.7
    f = np.zeros((R, ld))
    f[:, :3 * N] = W.forces(61, N, R)
    v0 = np.full((R, ld), 5.0)
    v0[:, :3 * N] = W.forces(62, N, R)
    ref = oracle.compute_spme(xi, K, p, L, pos, f[:, :3 * N], alpha=2.1, beta=2.2,
                              u=v0[:, :3 * N])
    d = E.DropIn(xi, K, p, N, L)
    v = v0.copy()
    d.compute(pos, R, 2.1, f, ld, 2.2, v, ld)
    assert _rel(v[:, :3 * N], ref) < FP64_TOL
    assert np.all(v[:, 3 * N:] == 5.0)   # padding untouched
    one = E.DropIn(xi, K, p, N, L)
    for r in range(R):
        vr = np.ascontiguousarray(v0[r])
        one.compute(pos if r == 0 else None, 1, 2.1, np.ascontiguousarray(f[r]), ld, 2.2, vr,
                    ld)
        np.testing.assert_array_equal(vr, v[r])
    d.close()
    one.close()


@pytest.mark.parametrize("p,K", [(4, 32), (6, 32), (8, 48)])
def test_spread_and_fft_stages(oracle, E, torch, p, K):
    _say(f"case: stages p={p} K={K} N=300")
    N = 300
    pos, L = _system(77 + p, N, 0.15)
    xi = 0.6
    f = W.forces(77 + p, N)[0]
    eng = E.Engine(xi, K, p, N, L, precision=E.SPME_FP64)
    eng.set_positions(pos)
    ft = torch.tensor(f, device="cuda").view(1, -1)
    m1 = eng.debug_stage(1, ft)[:, :, :, :K]
    P_, ind, _, _ = oracle.compute_P(K, p, pos, L)
    g = oracle.spread(K, p, P_, ind, f)
    assert _rel(m1, g) < 1e-13
    m2 = eng.debug_stage(2, ft)[:, :, :, :K]
    lm2 = oracle.compute_lm2(xi, K, p, L)
    spec = np.stack([oracle.fft3d_r2c(g[c]) for c in range(3)])
    spec = oracle.apply_influence(K, L, lm2, spec)
    back = np.stack([oracle.fft3d_c2r(spec[c]) for c in range(3)]) / L ** 3
    assert _rel(m2, back) < FP64_TOL
    eng.close()


@pytest.mark.parametrize("prec,K", [("fp64", 96), ("fp64", 128), ("fp64", 192),
                                    ("fp32", 128), ("fp32", 192), ("fp32", 256)])
def test_fft_pipeline_large_meshes(oracle, E, torch, prec, K):
    """Forward transform, influence and inverse transform on the engine's own
    spread mesh, at the sizes where whole planes are transformed in LDS (FP64
    K <= 128, FP32 K <= 192) and beyond (separate x and y passes).  The
    reference transforms are numpy's rfftn / irfftn (the c2r drops the
    imaginary parts of the kx = 0 and K/2 terms, as the engine's does) around
    the oracle's influence step."""
    _say(f"case: fft pipeline {prec} K={K} N=2000 p=6")
    N, p = 2000, 6
    pos, L = _system(300 + K, N, 0.2)
    xi = 0.45
    f = W.forces(300 + K, N)[0]
    P = E.SPME_FP64 if prec == "fp64" else E.SPME_FP32
    dt = torch.float64 if prec == "fp64" else torch.float32
    eng = E.Engine(xi, K, p, N, L, precision=P)
    eng.set_positions(pos)
    ft = torch.tensor(f, device="cuda", dtype=dt).view(1, -1)
    m1 = eng.debug_stage(1, ft)[:, :, :, :K].astype(np.float64)
    m2 = eng.debug_stage(2, ft)[:, :, :, :K]
    eng.close()
    lm2 = oracle.compute_lm2(xi, K, p, L)
    spec = np.stack([np.fft.rfftn(m1[c]) for c in range(3)])
    spec = oracle.apply_influence(K, L, lm2, spec)
    back = np.stack([np.fft.irfftn(spec[c], s=(K, K, K), axes=(0, 1, 2)) for c in range(3)]) * (K ** 3 / L ** 3)
    assert _rel(m2, back) < (FP64_TOL if prec == "fp64" else FP32_TOL)


@pytest.mark.parametrize("prec,K", [("fp64", 128), ("fp32", 192), ("fp32", 256)])
def test_fft_pipeline_vs_rocfft(oracle, E, torch, prec, K):
    """The same pipeline against rocFFT (torch.fft on the GPU) in the engine's
    precision: r2c over the whole mesh, the oracle's influence, then c2c
    along z and y and a c2r along x whose k_x = 0 and K/2 terms are made real
    first (the engine's convention, and a well-defined c2r input)."""
    _say(f"case: fft pipeline vs rocFFT {prec} K={K} N=2000 p=6")
    N, p = 2000, 6
    pos, L = _system(400 + K, N, 0.2)
    xi = 0.45
    f = W.forces(400 + K, N)[0]
    P = E.SPME_FP64 if prec == "fp64" else E.SPME_FP32
    dt = torch.float64 if prec == "fp64" else torch.float32
    eng = E.Engine(xi, K, p, N, L, precision=P)
    eng.set_positions(pos)
    ft = torch.tensor(f, device="cuda", dtype=dt).view(1, -1)
    m1 = eng.debug_stage(1, ft)[:, :, :, :K]
    m2 = eng.debug_stage(2, ft)[:, :, :, :K]
    eng.close()
    lm2 = oracle.compute_lm2(xi, K, p, L)
    spec = torch.fft.rfftn(torch.tensor(np.ascontiguousarray(m1), device="cuda"), dim=(1, 2, 3))
    spec = oracle.apply_influence(K, L, lm2, spec.cpu().numpy().astype(np.complex128))
    g = torch.tensor(spec, device="cuda").to(torch.complex128 if prec == "fp64" else torch.complex64)
    g = torch.fft.ifftn(g, dim=(1, 2))
    g[..., 0] = g[..., 0].real.to(g.dtype)
    g[..., K // 2] = g[..., K // 2].real.to(g.dtype)
    back = torch.fft.irfft(g, n=K, dim=3).double().cpu().numpy() * (K ** 3 / L ** 3)
    assert _rel(m2, back) < (FP64_TOL if prec == "fp64" else FP32_TOL)


def test_c2_fp64_fp32_vs_oracle_and_direct(oracle, E, torch):
    """The plain-vs-FP32 workload: N = 100,000, phi = 0.2, K = 128, p = 6,
    rc = 4, one right-hand side, the same xi (the one minimising the error
    estimate at that cutoff) for both precisions.  Full operator against the
    oracle's full M.f, and both against the committed direct-Ewald rows (at
    rc = 4 the real-space truncation dominates: the estimate is 1.3e-2)."""
    import os
    from spme_amd import params as PR
    _say("case: c2 N=100000 K=128 p=6 rc=4 fp64+fp32 vs oracle and direct")
    cfg = W.CONFIGS["c2"]
    N, K, p, rc = cfg["n"], cfg["K"], cfg["p"], cfg["rc"]
    sy = W.make_system(cfg["seed"], N, cfg["phi"])
    xi = 0.58
    f = W.forces(cfg["seed"], N, 1).astype(np.float32).astype(np.float64)
    ref = oracle.spme_full(xi, K, p, rc, sy.L, sy.pos, f)
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "c2.npz"))
    rows, direct = g["rows"], g["direct"][0]
    est = PR.estimate_error(xi, K, p, rc, cfg["phi"], sy.L)
    errs = {}
    for prec, tol, dt in [(E.SPME_FP64, FP64_TOL, torch.float64),
                          (E.SPME_FP32, FP32_TOL, torch.float32)]:
        eng = E.Engine(xi, K, p, N, sy.L, rc=rc, precision=prec)
        eng.set_positions(torch.tensor(sy.pos, device="cuda"))
        u = eng.apply(torch.tensor(f, device="cuda", dtype=dt), parts=E.SPME_FULL)
        got = u.double().cpu().numpy()
        eng.close()
        assert _rel(got, ref) < tol, prec
        mine = got[0].reshape(N, 3)[rows].reshape(-1)
        errs[prec] = float(np.linalg.norm(mine - direct) / np.linalg.norm(direct))
    _say(f"  error vs direct: fp64 {errs[E.SPME_FP64]:.3e}, fp32 {errs[E.SPME_FP32]:.3e}, "
         f"estimate {est:.3e}")
    assert abs(errs[E.SPME_FP32] - errs[E.SPME_FP64]) < 1e-4 * errs[E.SPME_FP64] + 1e-6
    assert est / 2 < errs[E.SPME_FP64] < 2 * est


@pytest.mark.parametrize("K,p", [(32, 4), (48, 6), (64, 8)])
def test_recip_fp64_and_fp32(oracle, E, torch, K, p):
    _say(f"case: recip K={K} p={p} N=1500 fp64+fp32")
    N = 1500
    pos, L = _system(5 + K, N, 0.2)
    xi = 0.55
    f = W.forces(5 + K, N, 1)
    ref = oracle.compute_spme(xi, K, p, L, pos, f)
    for prec, tol, dt in [(E.SPME_FP64, FP64_TOL, torch.float64),
                          (E.SPME_FP32, FP32_TOL, torch.float32)]:
        eng = E.Engine(xi, K, p, N, L, precision=prec)
        eng.set_positions(torch.tensor(pos, device="cuda"))
        u = eng.apply(torch.tensor(f, device="cuda", dtype=dt))
        torch.cuda.synchronize()
        assert _rel(u.double().cpu().numpy(), ref) < tol, prec
        eng.close()


@pytest.mark.parametrize("K,p,N", [(32, 4, 6000), (32, 6, 8000), (48, 8, 15000)])
def test_recip_dense_systems(oracle, E, torch, K, p, N):
    """0.18-0.24 particles per mesh cell: the spreading's staging groups and
    single steps exceed one batch, so steps are staged one by one in
    several batches (the path the sparse workloads rarely take)."""
    _say(f"case: recip dense K={K} p={p} N={N} fp64+fp32, 4 RHS")
    pos, L = _system(7 + p, N, 0.2)
    xi = 0.5
    f = W.forces(7 + p, N, 4)
    ref = oracle.compute_spme(xi, K, p, L, pos, f)
    for prec, tol, dt in [(E.SPME_FP64, FP64_TOL, torch.float64),
                          (E.SPME_FP32, FP32_TOL, torch.float32)]:
        eng = E.Engine(xi, K, p, N, L, precision=prec, max_rhs=4)
        eng.set_positions(torch.tensor(pos, device="cuda"))
        u = eng.apply(torch.tensor(f, device="cuda", dtype=dt))
        torch.cuda.synchronize()
        assert _rel(u.double().cpu().numpy(), ref) < tol, prec
        eng.close()


def test_batched_rhs_equal_single(oracle, E, torch):
    _say("case: batched R=16 fp32 and R=4 fp64, K=48 p=8 N=2000")
    N, K, p = 2000, 48, 8
    pos, L = _system(91, N, 0.2)
    xi = 0.5
    f = W.forces(91, N, 16)
    ref = oracle.compute_spme(xi, K, p, L, pos, f)
    eng = E.Engine(xi, K, p, N, L, precision=E.SPME_FP32, max_rhs=16)
    eng.set_positions(pos)
    u = eng.apply(torch.tensor(f, device="cuda", dtype=torch.float32))
    got = u.double().cpu().numpy()
    for r in range(16):
        assert _rel(got[r], ref[r]) < FP32_TOL, r
    eng.close()
    eng = E.Engine(xi, K, p, N, L, precision=E.SPME_FP64, max_rhs=4)
    eng.set_positions(pos)
    u = eng.apply(torch.tensor(f[:4], device="cuda"))
    single = E.Engine(xi, K, p, N, L, precision=E.SPME_FP64, max_rhs=1)
    single.set_positions(pos)
    for r in range(4):
        assert _rel(u[r].cpu().numpy(), ref[r]) < FP64_TOL
        s = single.apply(torch.tensor(f[r], device="cuda"))
        # same kernels and summation order: batching changes nothing
        np.testing.assert_array_equal(s[0].cpu().numpy(), u[r].cpu().numpy())
    eng.close()
    single.close()


def test_strided_alpha_beta(oracle, E, torch):
    _say("case: ldin/ldout > 3N, alpha/beta")
    N, K, p = 500, 32, 6
    pos, L = _system(13, N, 0.1)
    xi = 0.5
    f = W.forces(13, N, 3)
    u0 = W.forces(14, N, 3)
    ref = oracle.compute_spme(xi, K, p, L, pos, f, alpha=-0.7, beta=1.9, u=u0)
    eng = E.Engine(xi, K, p, N, L, max_rhs=3)
    eng.set_positions(pos)
    fin = torch.zeros(3, 3 * N + 16, dtype=torch.float64, device="cuda")
    fin[:, :3 * N] = torch.tensor(f)
    uo = torch.full((3, 3 * N + 40), 7.0, dtype=torch.float64, device="cuda")
    uo[:, :3 * N] = torch.tensor(u0)
    eng.apply(fin[:, :3 * N], uo[:, :3 * N], alpha=-0.7, beta=1.9)
    got = uo.cpu().numpy()
    assert _rel(got[:, :3 * N], ref) < FP64_TOL
    assert np.all(got[:, 3 * N:] == 7.0)   # padding untouched
    eng.close()


def test_full_mobility_vs_oracle_and_direct(oracle, E, torch):
    _say("case: full M.f fp64/fp32 vs oracle spme_full and direct Ewald")
    N = 1000
    pos, L = _system(1001, N, 0.1)
    # rc = 10 and xi = 0.5: real-space tail ~1e-9, xi*h = 0.27 at K = 64, p = 8
    rc, xi = 10.0, 0.5
    f = W.forces(1001, N, 2)
    ref = oracle.spme_full(xi, 64, 8, rc, L, pos, f)
    rows = W.rows(3, N, 200)
    direct = oracle.ewald_direct(pos, L, f[0], rows=rows)
    # the oracle's own SPME is 3.2e-6 from the direct sum at these parameters
    for prec, tol, dt, acc in [(E.SPME_FP64, FP64_TOL, torch.float64, 1e-5),
                               (E.SPME_FP32, FP32_TOL, torch.float32, 1e-5)]:
        eng = E.Engine(xi, 64, 8, N, L, rc=rc, precision=prec, max_rhs=2)
        eng.set_positions(pos)
        u = eng.apply(torch.tensor(f, device="cuda", dtype=dt), parts=E.SPME_FULL)
        got = u.double().cpu().numpy()
        assert _rel(got, ref) < tol, prec
        eng.close()
        mine = got[0].reshape(N, 3)[rows].reshape(-1)
        err = np.linalg.norm(mine - direct) / np.linalg.norm(direct)
        assert err < acc, (prec, err)


def test_fp32_sum64_form(oracle, E, torch):
    """SPME_FP32_SUM64: the FP32 mesh path whose final beta*u + alpha*(recip +
    real) sum is formed in FP64 (float inputs, double outputs).  Against the
    oracle at the FP32 bar, with alpha and beta; against the FP32 engine the
    only difference is the final rounding (within 2e-7 relative); and the
    outputs carry bits beyond FP32 (most entries), so the sum was not
    rounded to float."""
    _say("case: fp32 mesh path with an fp64 final sum, full M.f, alpha/beta")
    N, K, p, rc, xi = 1000, 64, 8, 10.0, 0.5
    pos, L = _system(1001, N, 0.1)
    f = W.forces(1001, N, 2).astype(np.float32).astype(np.float64)
    u0 = W.forces(1005, N, 2)
    ref = oracle.spme_full(xi, K, p, rc, L, pos, f)
    eng = E.Engine(xi, K, p, N, L, rc=rc, precision=E.SPME_FP32_SUM64, max_rhs=2)
    eng.set_positions(pos)
    f32 = torch.tensor(f, device="cuda", dtype=torch.float32)
    u = eng.apply(f32, parts=E.SPME_FULL)
    assert u.dtype == torch.float64
    got = u.cpu().numpy()
    assert _rel(got, ref) < FP32_TOL
    ub = torch.tensor(u0, device="cuda")
    eng.apply(f32, ub, alpha=-0.7, beta=1.9, parts=E.SPME_FULL)
    assert _rel(ub.cpu().numpy(), 1.9 * u0 - 0.7 * ref) < FP32_TOL
    # beta*u0 keeps FP64 precision: the result minus the FP64 reference part
    # is the FP32 operator's error only
    assert np.abs(ub.cpu().numpy() - (1.9 * u0 - 0.7 * got)).max() < 1e-12
    # the sum of the two FP32 parts is itself an FP32 value when their
    # exponents are close (70 % of entries here are not)
    assert np.mean(got != got.astype(np.float32).astype(np.float64)) > 0.5
    eng.close()
    e32 = E.Engine(xi, K, p, N, L, rc=rc, precision=E.SPME_FP32, max_rhs=2)
    e32.set_positions(pos)
    g32 = e32.apply(f32, parts=E.SPME_FULL).double().cpu().numpy()
    e32.close()
    assert _rel(g32, got) < 2e-7
    with pytest.raises(E.SpmeError, match="groups take"):
        g = E.Engine(xi, K, p, N, L, rc=rc, precision=E.SPME_FP32_SUM64, max_rhs=2)
        g.set_positions(pos)
        E.Group([g], 2)


def test_deterministic(E, torch):
    _say("case: run twice bitwise, fp32 K=48 p=6 R=4")
    N, K, p = 3000, 48, 6
    pos, L = _system(55, N, 0.25)
    f = torch.tensor(W.forces(55, N, 4), device="cuda", dtype=torch.float32)
    outs = []
    for _ in range(2):
        eng = E.Engine(0.6, K, p, N, L, rc=4.0, precision=E.SPME_FP32, max_rhs=4)
        eng.set_positions(pos)
        outs.append(eng.apply(f, parts=E.SPME_FULL).cpu().numpy())
        eng.close()
    np.testing.assert_array_equal(outs[0], outs[1])


def test_errors_are_reported(E, torch):
    _say("case: argument errors")
    with pytest.raises(E.SpmeError, match="porder"):
        E.Engine(0.5, 32, 5, 100, 10.0)
    with pytest.raises(E.SpmeError, match="dim"):
        E.Engine(0.5, 40, 4, 100, 10.0)
    with pytest.raises(E.SpmeError, match="4 \\* porder"):
        E.Engine(0.5, 32, 8, 100, 10.0)
    eng = E.Engine(0.5, 32, 6, 10, 10.0, max_rhs=2)
    f = torch.zeros(3, 30, dtype=torch.float64, device="cuda")
    with pytest.raises(E.SpmeError, match="positions"):
        eng.apply(f[:1])
    bad = np.zeros((10, 3))
    bad[3, 1] = np.nan
    with pytest.raises(E.SpmeError, match="finite"):
        eng.set_positions(bad)
    eng.set_positions(np.full((10, 3), 1.0))
    with pytest.raises(E.SpmeError, match="nrhs"):
        eng.apply(f)
    eng.close()
