"""Training data for the learned parameter selector: seeded SPME runs on the
GPU, each scored by its measured relative error against the FP64 direct
Beenakker-Ewald sum (the CPU oracle, on a 64-row subset) and by its
measured apply time.

Systems: N log-uniform in [1e3, 1e5], phi uniform in [0.05, 0.35].
Per system, 20 parameter sets: the analytic selection for tolerances
1e-3 .. 1e-7 and three random perturbations of each (neighbouring mesh,
other B-spline order, xi scaled by U[0.8, 1.25], rc from the real-space
budget or scaled by U[0.85, 1.15]).  Errors are of the full M.f in FP64
(one right-hand side); times are FP64 single-RHS applies.

    python tests/golden/make_mlp_data.py START COUNT OUT.npz
"""
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from spme_amd import engine as E  # noqa: E402
from spme_amd import params as PR  # noqa: E402
from spme_amd import workload as W  # noqa: E402

TOLS = (1e-3, 1e-4, 1e-5, 1e-6, 1e-7)
NROWS = 64


def system(s):
    u = W.uniform(90000 + s, 3, 2)
    n = int(round(math.exp(math.log(1e3) + u[0] * (math.log(1e5) - math.log(1e3)))))
    phi = 0.05 + 0.30 * float(u[1])
    return n, phi


def candidates(n, phi, rng):
    L = PR.box_length(n, phi)
    out = []
    for tol in TOLS:
        try:
            a = PR.analytic_select(n, phi, tol)
        except ValueError:
            continue
        out.append((a.xi, a.K, a.p, a.rc))
        ks = list(PR.SMOOTH_K)
        for _ in range(3):
            ki = ks.index(a.K) + int(rng.integers(-1, 2))
            K = ks[min(max(ki, 0), len(ks) - 1)]
            p = int(rng.choice([4, 6, 8]))
            while 4 * p >= K:
                p -= 2
            xi = a.xi * float(rng.uniform(0.8, 1.25))
            rc = PR.rc_for(xi, tol * PR.rms_u(phi, L) * 0.7, phi, L / 3.0)
            if rc is None:
                rc = min(a.rc * float(rng.uniform(0.85, 1.15)), L / 3.0)
            out.append((xi, K, p, rc))
    return out


def main(start, count, path):
    import torch
    rec = {k: [] for k in ("sys", "n", "phi", "L", "xi", "K", "p", "rc", "err", "time")}
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_all = time.time()
    for s in range(start, start + count):
        n, phi = system(s)
        sy = W.make_system(70000 + s, n, phi)
        f = W.forces(70000 + s, n, 1)
        rows = W.rows(70000 + s, n, NROWS)
        t0 = time.time()
        direct = O.ewald_direct(sy.pos, sy.L, f[0], rows=rows)
        t_direct = time.time() - t0
        rng = np.random.default_rng(s)
        pos_d = torch.tensor(sy.pos, device="cuda")
        f_d = torch.tensor(f, device="cuda")
        for (xi, K, p, rc) in candidates(n, phi, rng):
            eng = E.Engine(xi, K, p, n, sy.L, rc=rc, precision=E.SPME_FP64, max_rhs=1)
            eng.set_positions(pos_d)
            u = eng.apply(f_d, parts=E.SPME_FULL)
            torch.cuda.synchronize()
            mine = u[0].cpu().numpy().reshape(n, 3)[rows].reshape(-1)
            err = float(np.linalg.norm(mine - direct) / np.linalg.norm(direct))
            ev0.record()
            for _ in range(3):
                eng.apply(f_d, u, parts=E.SPME_FULL)
            ev1.record()
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1) / 3
            eng.close()
            for k, v in zip(rec, (s, n, phi, sy.L, xi, K, p, rc, err, ms * 1e-3)):
                rec[k].append(v)
        print(f"system {s}: N={n} phi={phi:.3f} direct {t_direct:.1f}s, "
              f"{len(rec['sys'])} runs, {time.time() - t_all:.0f}s", flush=True)
    np.savez_compressed(path, **{k: np.array(v) for k, v in rec.items()})
    print("wrote", path, flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]), int(sys.argv[2]), sys.argv[3])
