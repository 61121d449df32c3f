"""The uneven multi-system workload through the grouped path, against the
oracle (tests/golden/c4.npz, written by tests/golden/make_c4.py).

64 periodic boxes (N log-uniform in [2e3, 5e4], phi in [0.05, 0.3], K in
{32, 48, 64, 96}, p = 6, 4 right-hand sides each, full M.f) applied as one
group: one launch per stage over every system, eagerly and as a captured
hipGraph replayed three times with eager work and synchronisation between
replays (each replay checked).  Every member's 200 stored rows must equal
the restated reference SPME + real space (oracle.spme_full):
  FP64 <= 1e-12, FP32 <= 2e-5 relative max-norm per member.
The reference applies one system and one right-hand side at a time
(old-version/spme.c:631-702); the group replaces that loop.

The CPU test checks that the fixture still describes workload.multi_systems()
(sizes, meshes, xi and rc), so a change of the workload or of the parameter
estimate cannot leave the GPU test checking a stale fixture.
"""
import os
import sys

import numpy as np
import pytest

from spme_amd import workload as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c4.npz")
FP64_TOL, FP32_TOL = 1e-12, 2e-5


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fixture():
    g = np.load(GOLD)
    return g["params"], g["rows"], g["full"]


def test_fixture_matches_workload():
    params, rows, full = _fixture()
    cfg = W.CONFIGS["c4"]
    ms = W.multi_systems(seed=cfg["seed"], nsys=cfg["nsys"], p=cfg["p"], nrhs=cfg["nrhs"])
    assert len(ms) == params.shape[0] == 64
    now = np.array([[m.seed, m.n, m.phi, m.L, m.K, m.p, m.xi, m.rc] for m in ms])
    np.testing.assert_allclose(now, params, rtol=1e-12, atol=0)
    assert {int(k) for k in params[:, 4]} == {32, 48, 64, 96}
    assert params[:, 1].min() >= 2000 and params[:, 1].max() <= 50000
    for s, m in enumerate(ms):
        np.testing.assert_array_equal(rows[s], W.rows(m.seed, m.n, rows.shape[1]))
    assert full.shape == (64, cfg["nrhs"], 3 * rows.shape[1])
    assert np.all(np.isfinite(full))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_c4_group_vs_oracle(prec):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from spme_amd import engine as E
    params, rows, full = _fixture()
    R = W.CONFIGS["c4"]["nrhs"]
    P = E.SPME_FP64 if prec == "fp64" else E.SPME_FP32
    dt = torch.float64 if prec == "fp64" else torch.float32
    tol = FP64_TOL if prec == "fp64" else FP32_TOL
    _say(f"case: c4, 64 uneven systems, {prec}, grouped full M.f (eager and graph)")
    engines, fs, us, idx = [], [], [], []
    for (seed, n, phi, L, K, p, xi, rc), r in zip(params, rows):
        seed, n, K, p = int(seed), int(n), int(K), int(p)
        e = E.Engine(xi, K, p, n, L, rc=rc, precision=P, max_rhs=R)
        e.set_positions(W.positions(seed, n, L))
        engines.append(e)
        f = W.forces(seed, n, R).astype(np.float32).astype(np.float64)
        fs.append(torch.tensor(f, device="cuda", dtype=dt))
        us.append(torch.zeros_like(fs[-1]))
        idx.append(torch.tensor((3 * r[:, None] + np.arange(3)[None, :]).reshape(-1),
                                device="cuda"))
    g = E.Group(engines, R)

    def check(what):
        torch.cuda.synchronize()
        worst = 0.0
        for s in range(len(engines)):
            got = us[s][:, idx[s]].double().cpu().numpy()
            d = float(np.abs(got - full[s]).max() / np.abs(full[s]).max())
            worst = max(worst, d)
            assert d < tol, (what, s, d)
        _say(f"  {what}: worst member {worst:.2e}")

    g.apply(fs, us, parts=E.SPME_FULL, graph=False)
    check("eager")
    scratch = torch.randn(1 << 20, device="cuda")
    for k in range(3):
        for u in us:
            u.fill_(float("nan"))
        g.apply(fs, us, parts=E.SPME_FULL, graph=True)
        check(f"graph replay {k}")
        scratch.mul_(1.0001)  # eager work between replays
    g.close()
    for e in engines:
        e.close()
