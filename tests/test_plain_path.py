"""The plain FP64 path (SPME_FP64_PLAIN): the reference's data structures
and call pattern (dense p^3 weights and int32 offsets per particle, atomic
spreading, one right-hand side at a time, influence and 1/K^3 scale passes)
on the GPU, through the C ABI.

Bars: <= 1e-12 relative max norm against the oracle's restatement of the
reference (old-version/spme.c:560-726) and against the compact FP64 path
(same operator, other summation order; the atomic spreading also makes the
plain path's last bits run-dependent).
"""
import sys

import numpy as np
import pytest

from spme_amd import workload as W

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


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


@pytest.mark.parametrize("K,p", [(32, 4), (48, 6), (64, 8)])
def test_plain_recip_vs_oracle(oracle, E, torch, K, p):
    _say(f"case: plain recip K={K} p={p} N=1500")
    N = 1500
    s = W.make_system(5 + K, N, 0.2)
    f = W.forces(5 + K, N, 1)
    ref = oracle.compute_spme(0.55, K, p, s.L, s.pos, f)
    eng = E.Engine(0.55, K, p, N, s.L, precision=E.SPME_FP64_PLAIN)
    assert eng.info().precision == E.SPME_FP64_PLAIN
    eng.set_positions(torch.tensor(s.pos, device="cuda"))
    u = eng.apply(torch.tensor(f, device="cuda"))
    torch.cuda.synchronize()
    assert _rel(u.cpu().numpy(), ref) < FP64_TOL
    eng.close()


def test_plain_rhs_loop_strided_alpha_beta(oracle, E, torch):
    _say("case: plain nrhs=3, ldin/ldout > 3N, alpha/beta, K=32 p=6 N=500")
    N, K, p = 500, 32, 6
    s = W.make_system(13, N, 0.1)
    f = W.forces(13, N, 3)
    u0 = W.forces(14, N, 3)
    ref = oracle.compute_spme(0.5, K, p, s.L, s.pos, f, alpha=-0.7, beta=1.9, u=u0)
    eng = E.Engine(0.5, K, p, N, s.L, precision=E.SPME_FP64_PLAIN, max_rhs=3)
    eng.set_positions(s.pos)
    fin = torch.zeros(3, 3 * N + 16, dtype=torch.float64, device="cuda")
    fin[:, :3 * N] = torch.tensor(f)
    uo = torch.full((3, 3 * N + 40), 7.0, dtype=torch.float64, device="cuda")
    uo[:, :3 * N] = torch.tensor(u0)
    eng.apply(fin[:, :3 * N], uo[:, :3 * N], alpha=-0.7, beta=1.9)
    got = uo.cpu().numpy()
    assert _rel(got[:, :3 * N], ref) < FP64_TOL
    assert np.all(got[:, 3 * N:] == 7.0)   # padding untouched
    # beta = 0 does not read the output (NaN there stays out of the result)
    un = torch.full((3, 3 * N), float("nan"), dtype=torch.float64, device="cuda")
    eng.apply(fin[:, :3 * N], un, alpha=1.0, beta=0.0)
    ref0 = oracle.compute_spme(0.5, K, p, s.L, s.pos, f)
    assert _rel(un.cpu().numpy(), ref0) < FP64_TOL
    eng.close()


@pytest.mark.parametrize("K,p,n", [(128, 6, 20000), (192, 8, 30000)])
def test_plain_equals_compact_large_meshes(E, torch, K, p, n):
    """Large meshes: the plain path's separate x, y, z passes and influence
    pass against the compact path's fused passes."""
    _say(f"case: plain vs compact fp64 recip K={K} p={p} N={n}")
    s = W.make_system(300 + K, n, 0.2)
    f = torch.tensor(W.forces(300 + K, n, 2), device="cuda")
    outs = []
    for prec in (E.SPME_FP64_PLAIN, E.SPME_FP64):
        eng = E.Engine(0.5, K, p, n, s.L, precision=prec, max_rhs=2)
        eng.set_positions(s.pos)
        outs.append(eng.apply(f).cpu().numpy())
        eng.close()
    assert _rel(outs[0], outs[1]) < FP64_TOL


def test_plain_full_operator_equals_compact(E, torch):
    _say("case: plain vs compact fp64, full M.f, K=48 p=6 N=3000 rc=4")
    N, K, p = 3000, 48, 6
    s = W.make_system(55, N, 0.25)
    f = torch.tensor(W.forces(55, N, 2), device="cuda")
    outs = []
    for prec in (E.SPME_FP64_PLAIN, E.SPME_FP64):
        eng = E.Engine(0.6, K, p, N, s.L, rc=4.0, precision=prec, max_rhs=2)
        eng.set_positions(s.pos)
        outs.append(eng.apply(f, parts=E.SPME_FULL).cpu().numpy())
        eng.close()
    assert _rel(outs[0], outs[1]) < FP64_TOL


def test_plain_rejects_tiles_groups_and_debug(E, torch):
    _say("case: plain-path error paths")
    N, K, p = 400, 32, 4
    s = W.make_system(21, N, 0.1)
    eng = E.Engine(0.5, K, p, N, s.L, precision=E.SPME_FP64_PLAIN)
    eng.set_positions(s.pos)
    with pytest.raises(E.SpmeError, match="plain path"):
        eng.set_tiles({"spread_cc": 4})
    with pytest.raises(E.SpmeError, match="plain path"):
        eng.autotune(nrhs=1, reps=1)
    with pytest.raises(E.SpmeError, match="cannot be grouped"):
        E.Group([eng], 1)
    eng.close()
