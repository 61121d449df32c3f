"""Write tests/golden/c4.npz: oracle rows of the uneven multi-system workload.

The 64 periodic boxes of workload.multi_systems() (N log-uniform in
[2e3, 5e4], phi in [0.05, 0.3], K in {32, 48, 64, 96}, p = 6, xi and rc
from the analytic estimate for 1e-4), 4 right-hand sides each, forces
rounded to FP32 so that FP32 and FP64 engines see the same inputs.  For each
system the restated reference SPME plus real space, overlap and self
(oracle.spme_full, the full M.f) is stored on 200 seeded rows, with the
system's parameters, so that the GPU test needs no oracle time.  The reference
applies systems and right-hand sides one at a time (old-version/spme.c:
631-702); the grouped path replaces that loop, and these rows are what each
member's result must equal.

    python tests/golden/make_c4.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from spme_amd import workload as W  # noqa: E402

NROWS = 200


def main():
    cfg = W.CONFIGS["c4"]
    R = cfg["nrhs"]
    members = W.multi_systems(seed=cfg["seed"], nsys=cfg["nsys"], p=cfg["p"], nrhs=R)
    params = np.array([[m.seed, m.n, m.phi, m.L, m.K, m.p, m.xi, m.rc] for m in members])
    rows, full = [], []
    t = time.time()
    for m in members:
        pos = W.positions(m.seed, m.n, m.L)
        f = W.forces(m.seed, m.n, R).astype(np.float32).astype(np.float64)
        r = W.rows(m.seed, m.n, NROWS)
        u = O.spme_full(m.xi, m.K, m.p, m.rc, m.L, pos, f)
        idx = (3 * r[:, None] + np.arange(3)[None, :]).reshape(-1)
        rows.append(r)
        full.append(u[:, idx])
    secs = time.time() - t
    np.savez_compressed(os.path.join(HERE, "c4.npz"), params=params, rows=np.stack(rows),
                        full=np.stack(full))
    meta = dict(config="c4", nsys=len(members), nrhs=R, rows_per_system=NROWS,
                forces="fp32-rounded", total_n=int(sum(m.n for m in members)),
                params_columns=["seed", "n", "phi", "L", "K", "p", "xi", "rc"],
                oracle_seconds=round(secs, 1), threads=O.num_threads())
    with open(os.path.join(HERE, "c4.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    print(meta)


if __name__ == "__main__":
    main()
