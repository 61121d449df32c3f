"""Held-out evaluation set for the learned parameter selector: systems not
used in training (other seeds), plus the large evaluation case N = 500,000,
phi = 0.3, each with the FP64 direct Beenakker-Ewald sum on 64 rows for one
seeded force vector (the CPU oracle).

    python tests/golden/make_mlp_eval.py
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
from spme_amd import workload as W  # noqa: E402

NROWS = 64


def eval_systems():
    out = []
    for s in range(16):
        u = W.uniform(40000 + s, 3, 2)
        n = int(round(math.exp(math.log(1e3) + u[0] * (math.log(1e5) - math.log(1e3)))))
        out.append((60000 + s, n, 0.05 + 0.30 * float(u[1])))
    out.append((1005, 500_000, 0.3))
    return out


def main():
    rec = {k: [] for k in ("seed", "n", "phi", "L")}
    rows_all, direct_all = [], []
    for seed, n, phi in eval_systems():
        t = time.time()
        s = W.make_system(seed, n, phi)
        f = W.forces(seed, n, 1)[0]
        rows = W.rows(seed, n, NROWS)
        d = O.ewald_direct(s.pos, s.L, f, rows=rows)
        for k, v in zip(rec, (seed, n, phi, s.L)):
            rec[k].append(v)
        rows_all.append(rows)
        direct_all.append(d)
        print(f"seed {seed} N={n} phi={phi:.3f}: {time.time() - t:.1f}s", flush=True)
    np.savez_compressed(os.path.join(HERE, "mlp_eval.npz"),
                        rows=np.stack(rows_all), direct=np.stack(direct_all),
                        **{k: np.array(v) for k, v in rec.items()})


if __name__ == "__main__":
    main()
