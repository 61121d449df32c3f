"""What the learned parameter selector gains over the analytic estimate.

Held-out systems (tests/golden/mlp_eval.npz: seeds not used in training,
plus N = 500,000 at phi = 0.3, outside the training range) with FP64
direct-Ewald rows.  For each system and target tolerance three selections
pick (xi, K, p, rc):
  analytic       the error model at the requested tolerance,
  analytic/4     the error model asked for tol/4 (the safety factor the
                 model needs to meet every held-out target),
  learned        the network; outside its training box (N = 500,000) its
                 inputs are clamped to the box and its scale-free outputs
                 (xi, mesh spacing, p, rc) carried over.
Each is run as the full M.f in FP64 with one right-hand side, the form the
training labels were measured in: the relative error against the direct rows
and the apply time.  The test checks that the learned selector meets the
tolerance at least as often as the plain analytic one and as the analytic
one with its safety factor (analytic/4), that it is faster overall than
analytic/4, and that at
N = 500,000 it acts (no fallback) and meets both tolerances, also in that
case's evaluation form: 8 right-hand sides batched, FP64 (and FP32 at 1e-4).
"""
import os

import numpy as np
import pytest

from spme_amd import workload as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOLS = (1e-4, 1e-6)
SAFETY = 4.0


def _run(torch, E, sy, f_d, rows, direct, prm, precision=None):
    """Error of right-hand side 0 against the direct rows, and the time of
    one apply of all rows of f_d."""
    xi, K, p, rc = prm
    n = sy.pos.shape[0]
    prec = E.SPME_FP64 if precision is None else precision
    eng = E.Engine(xi, K, p, n, sy.L, rc=rc, precision=prec, max_rhs=f_d.shape[0])
    eng.set_positions(sy.pos)
    u = eng.apply(f_d, parts=E.SPME_FULL)
    mine = u[0].double().cpu().numpy().reshape(n, 3)[rows].reshape(-1)
    err = float(np.linalg.norm(mine - direct) / np.linalg.norm(direct))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        eng.apply(f_d, u, parts=E.SPME_FULL)
    e1.record()
    torch.cuda.synchronize()
    eng.close()
    return err, e0.elapsed_time(e1) / 3


def test_learned_selector_gain():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from spme_amd import engine as E
    from spme_amd import mlp
    from spme_amd import params as PR
    if not os.path.exists(mlp.WEIGHTS):
        pytest.skip("no trained weights")
    net = mlp.Net.load()
    ev = np.load(os.path.join(HERE, "golden", "mlp_eval.npz"))
    rows_all, direct_all = ev["rows"], ev["direct"]
    res = {"analytic": [], "analytic/4": [], "learned": []}
    large = []  # (how, met) of the learned picks at N = 500,000
    for k in range(len(ev["seed"])):
        seed, n, phi = int(ev["seed"][k]), int(ev["n"][k]), float(ev["phi"][k])
        sy = W.make_system(seed, n, phi)
        f_d = torch.tensor(W.forces(seed, n, 1), device="cuda")
        for tol in TOLS:
            picks = {"analytic": PR.analytic_select(n, phi, tol).as_tuple(),
                     "analytic/4": PR.analytic_select(n, phi, tol / SAFETY).as_tuple()}
            xi, K, p, rc, how = net.select(n, phi, tol)
            picks["learned"] = (xi, K, p, rc)
            line = f"N={n:7d} phi={phi:.3f} tol={tol:.0e}"
            for name, prm in picks.items():
                err, ms = _run(torch, E, sy, f_d, rows_all[k], direct_all[k], prm)
                res[name].append((err <= tol, ms))
                if name == "learned" and n >= 500_000:
                    large.append((how, err <= tol))
                    # the evaluation form of the large case: 8 right-hand
                    # sides batched (row 0 is the force the direct rows
                    # belong to), FP64, and FP32 where it can reach the tolerance
                    f8 = W.forces(seed, n, 8)
                    assert np.array_equal(f8[0], W.forces(seed, n, 1)[0])
                    forms = [("fp64", E.SPME_FP64, torch.float64)]
                    if tol >= 1e-5:
                        forms.append(("fp32", E.SPME_FP32, torch.float32))
                    for tag, P, dt in forms:
                        e8, ms8 = _run(torch, E, sy, torch.tensor(f8, device="cuda", dtype=dt),
                                       rows_all[k], direct_all[k], prm, precision=P)
                        print(f"  N={n} tol={tol:.0e} learned, 8 RHS {tag}: err={e8:.2e} "
                              f"{ms8:.3f} ms ({8e3 / ms8:.0f} M.f/s)", flush=True)
                        large.append((how, e8 <= tol))
                tag = f"{name}({how})" if name == "learned" else name
                line += (f" | {tag} K={prm[1]} p={prm[2]} xi={prm[0]:.3f} rc={prm[3]:.2f}"
                         f" err={err:.2e} {ms:.3f}ms")
            print(line, flush=True)
    summ = {k: (sum(ok for ok, _ in v), sum(ms for _, ms in v)) for k, v in res.items()}
    total = len(res["learned"])
    for k, (ok, ms) in summ.items():
        print(f"{k:11s} meets tolerance {ok}/{total}, total time {ms:.2f} ms", flush=True)
    assert len(large) == 2 * len(TOLS) + 1 and all(h == "learned" and ok for h, ok in large), large
    assert summ["learned"][0] >= summ["analytic"][0]
    assert summ["learned"][0] >= summ["analytic/4"][0]
    assert summ["learned"][1] < summ["analytic/4"][1]
