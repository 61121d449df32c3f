"""Systems and direct-Ewald rows for the fixed-mesh parameter selector.

The fixed-mesh selector chooses (xi, rc) when the mesh K and the B-spline
order p are given (the batched workload fixes K = 192, p = 8).  Its training
runs (tests/golden/make_mesh_runs.py, on the GPU) are scored against the FP64
direct Beenakker-Ewald sum of the CPU oracle on 64 seeded rows; this script
writes those rows, so the GPU run needs no oracle time.

Systems: N log-uniform in [3e3, 3e4], phi uniform in [0.05, 0.35], one seeded
N(0,1) force vector each.

    python tests/golden/make_mesh_systems.py
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

NSYS, NROWS = 40, 64


def systems():
    out = []
    for s in range(NSYS):
        u = W.uniform(81000 + s, 3, 2)
        n = int(round(math.exp(math.log(3e3) + u[0] * (math.log(3e4) - math.log(3e3)))))
        out.append((82000 + s, n, 0.05 + 0.30 * float(u[1])))
    return out


def main():
    rec = {k: [] for k in ("seed", "n", "phi", "L")}
    rows_all, direct_all = [], []
    t0 = time.time()
    for seed, n, phi in systems():
        s = W.make_system(seed, n, phi)
        f = W.forces(seed, n, 1)[0]
        rows = W.rows(seed, n, NROWS)
        rows_all.append(rows)
        direct_all.append(O.ewald_direct(s.pos, s.L, f, rows=rows))
        for k, v in zip(rec, (seed, n, phi, s.L)):
            rec[k].append(v)
    np.savez_compressed(os.path.join(HERE, "mesh_direct.npz"), rows=np.stack(rows_all),
                        direct=np.stack(direct_all), **{k: np.array(v) for k, v in rec.items()})
    print(f"{NSYS} systems, {time.time() - t0:.0f}s", flush=True)


if __name__ == "__main__":
    main()
