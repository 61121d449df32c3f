"""Unequal radii on the GPU against the CPU oracle (spme_full_radii).

Engines created with radii carry f and a^2 f through the mesh path (six
components per right-hand side), apply the influence of the pair
psi_ij = A(k)(1 - k^2 (a_i^2 + a_j^2)/6) in the z pass, interpolate both
outputs and combine them per particle, and use the pair's radii in the
real-space coefficients (Ewald split of the unequal-sphere RPY tensor, Zuk
et al. 2014 overlap) and a per-particle self term.  Bars: FP64 <= 1e-12,
FP32 <= 2e-5 relative max-norm, as for unit radii.
"""
import sys

import numpy as np
import pytest

from spme_amd import workload as W

pytestmark = pytest.mark.gpu

FP64_TOL, FP32_TOL = 1e-12, 2e-5


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


def _system(seed, n, phi, amin=0.5, amax=1.8):
    L = W.box_length(n, phi)
    pos = W.positions(seed, n, L)
    rad = amin + (amax - amin) * W.uniform(seed, 4, n)
    return pos, rad, L


@pytest.mark.parametrize("K,p", [(32, 6), (48, 8)])
def test_radii_full_operator_vs_oracle(oracle, E, torch, K, p):
    N, xi, rc = 700, 0.6, 6.5
    pos, rad, L = _system(31, N, 0.15)
    f = W.forces(31, N, 3).astype(np.float32).astype(np.float64)
    u0 = W.forces(32, N, 3)
    _say(f"case: radii in [0.5, 1.8], N={N} L={L:.1f} K={K} p={p}, full and recip, alpha/beta")
    ref = oracle.spme_full_radii(xi, K, p, rc, L, pos, rad, f)
    refab = oracle.spme_full_radii(xi, K, p, rc, L, pos, rad, f, alpha=-0.7, beta=1.9, u=u0)
    for prec, tol, dt, dto in [(E.SPME_FP64, FP64_TOL, torch.float64, torch.float64),
                               (E.SPME_FP32, FP32_TOL, torch.float32, torch.float32),
                               (E.SPME_FP32_SUM64, FP32_TOL, torch.float32, torch.float64)]:
        eng = E.Engine(xi, K, p, N, L, rc=rc, precision=prec, max_rhs=3, radii=True)
        eng.set_radii(rad)
        eng.set_positions(pos)
        fd = torch.tensor(f, device="cuda", dtype=dt)
        got = eng.apply(fd, parts=E.SPME_FULL).double().cpu().numpy()
        d = _rel(got, ref)
        _say(f"  prec {prec}: full {d:.2e}")
        assert d < tol, (prec, d)
        ub = torch.tensor(u0, device="cuda", dtype=dto)
        eng.apply(fd, ub, alpha=-0.7, beta=1.9, parts=E.SPME_FULL)
        assert _rel(ub.double().cpu().numpy(), refab) < tol, prec
        # parts separately add up: reciprocal alone, then real space and self
        ur = eng.apply(fd, parts=E.SPME_RECIP)
        eng.apply(fd, ur, beta=1.0, parts=E.SPME_REAL | E.SPME_SELF)
        assert _rel(ur.double().cpu().numpy(), ref) < tol, prec
        eng.close()


def test_radii_single_rhs_matches_batch_and_direct(oracle, E, torch):
    """One right-hand side at a time equals the batch bitwise (FP64); the
    engine's M.f converges to the direct Ewald sum with radii."""
    N, xi, rc, K, p = 500, 0.5, 8.5, 64, 8  # the oracle: 2.3e-6 from the direct sum
    pos, rad, L = _system(33, N, 0.1, 0.7, 2.0)
    f = W.forces(33, N, 4)
    eng = E.Engine(xi, K, p, N, L, rc=rc, precision=E.SPME_FP64, max_rhs=4, radii=True)
    eng.set_radii(torch.tensor(rad, device="cuda"))
    eng.set_positions(torch.tensor(pos, device="cuda"))
    fd = torch.tensor(f, device="cuda")
    ub = eng.apply(fd, parts=E.SPME_FULL)
    for r in range(4):
        ur = eng.apply(fd[r:r + 1], parts=E.SPME_FULL)
        assert torch.equal(ur[0], ub[r]), r
    rows = W.rows(33, N, 100)
    d = oracle.ewald_direct_radii(pos, rad, L, f[0], rows=rows)
    mine = ub[0].cpu().numpy().reshape(N, 3)[rows].reshape(-1)
    err = np.linalg.norm(mine - d) / np.linalg.norm(d)
    _say(f"case: radii in [0.7, 2.0], vs direct Ewald: {err:.2e}")
    assert err < 1e-5
    eng.close()


def test_unit_radii_equal_the_plain_engine(E, torch):
    N, xi, rc, K, p = 800, 0.6, 6.0, 32, 6
    pos, _, L = _system(35, N, 0.2)
    f = torch.tensor(W.forces(35, N, 2), device="cuda")
    a = E.Engine(xi, K, p, N, L, rc=rc, precision=E.SPME_FP64, max_rhs=2)
    a.set_positions(pos)
    b = E.Engine(xi, K, p, N, L, rc=rc, precision=E.SPME_FP64, max_rhs=2, radii=True)
    b.set_positions(pos)  # radii default to 1
    ua = a.apply(f, parts=E.SPME_FULL).cpu().numpy()
    ub = b.apply(f, parts=E.SPME_FULL).cpu().numpy()
    assert _rel(ub, ua) < 1e-13
    a.close()
    b.close()


def test_radii_errors(E, torch):
    N = 200
    pos, rad, L = _system(37, N, 0.1)
    plain = E.Engine(0.5, 32, 6, N, L, rc=5.0)
    with pytest.raises(E.SpmeError, match="created without radii"):
        plain.set_radii(rad)
    plain.close()
    with pytest.raises(E.SpmeError, match="max_rhs"):
        E.Engine(0.5, 32, 6, N, L, rc=5.0, max_rhs=9, radii=True)
    with pytest.raises(E.SpmeError, match="plain path"):
        E.Engine(0.5, 32, 6, N, L, rc=5.0, precision=E.SPME_FP64_PLAIN, radii=True)
    eng = E.Engine(0.5, 32, 6, N, L, rc=5.0, max_rhs=2, radii=True)
    eng.set_positions(pos)
    bad = rad.copy()
    bad[3] = -1.0
    with pytest.raises(E.SpmeError, match="positive"):
        eng.set_radii(bad)
    f = torch.zeros(1, 3 * N, dtype=torch.float64, device="cuda")
    with pytest.raises(E.SpmeError, match="positions have not been set"):
        eng.apply(f)
    with pytest.raises(E.SpmeError, match="box_size"):
        eng.set_radii(np.full(N, L / 5.0))  # 2 a_max > L/3: no cell list
    eng.set_radii(rad)
    eng.set_positions(pos)
    eng.apply(f, parts=E.SPME_FULL)
    with pytest.raises(E.SpmeError, match="cannot be grouped"):
        E.Group([eng], 1)
    eng.close()
