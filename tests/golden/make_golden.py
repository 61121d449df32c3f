"""Write the golden fixtures under tests/golden/ from the CPU oracle.

The reference ships no fixtures (its verification dumps are commented out,
old-version/spme.c:623-718), so the fixtures are oracle outputs on the
seeded workloads: the restated reference SPME (recip part) and the FP64
direct Beenakker-Ewald sum, which does not depend on any SPME parameter.

    python tests/golden/make_golden.py [c1] [c2] [c3]
"""
import hashlib
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


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def make(name):
    cfg = W.CONFIGS[name]
    s = W.make_system(cfg["seed"], cfg["n"], cfg["phi"])
    nr = 2 if cfg["nrhs"] > 1 else 1
    # forces rounded to FP32 so that FP32 and FP64 runs see the same input
    f = W.forces(cfg["seed"], cfg["n"], cfg["nrhs"])[:nr].astype(np.float32)
    f = f.astype(np.float64)
    rows = (np.arange(cfg["n"], dtype=np.int32) if cfg["n"] <= 2000
            else W.rows(cfg["seed"], cfg["n"], 2000))
    t = time.time()
    direct = np.stack([O.ewald_direct(s.pos, s.L, f[r], rows=rows)
                       for r in range(nr)])
    t_direct = time.time() - t
    out = dict(rows=rows, direct=direct)
    meta = dict(config=name, n=cfg["n"], phi=cfg["phi"], L=s.L,
                seed=cfg["seed"], nrhs_stored=nr, forces="fp32-rounded",
                pos_sha=digest(s.pos), f_sha=digest(f),
                direct_xi=O.default_direct_xi(cfg["n"], s.L, rows.size),
                direct_seconds=round(t_direct, 1), threads=O.num_threads())
    if name == "c1":
        # restated reference SPME at the configuration-1 parameters
        xi = 0.8962270413934827  # cutoff2xi(5, 1e-6): |M_real(rc)| = 1e-6
        meta.update(K=cfg["K"], p=cfg["p"], rc=cfg["rc"], xi=xi)
        out["recip"] = O.compute_spme(xi, cfg["K"], cfg["p"], s.L, s.pos, f)
        out["full"] = O.spme_full(xi, cfg["K"], cfg["p"], cfg["rc"], s.L, s.pos, f)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    with open(os.path.join(HERE, f"{name}.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    print(name, meta)


if __name__ == "__main__":
    for n in sys.argv[1:] or ["c1", "c2", "c3"]:
        make(n)
