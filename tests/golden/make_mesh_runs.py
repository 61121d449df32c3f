"""Training runs of the fixed-mesh parameter selector, on the GPU.

For each system of tests/golden/mesh_direct.npz and each mesh (up to three
K with h = L/K in [0.45, 1.35], every order p in {4, 6, 8} with 4p < K), the
engine's full M.f in FP64 is scored against the direct-Ewald rows over a grid
of splitting parameters xi (xi h in [0.1, 0.8]) and cutoffs rc = t/xi
(t in [1.5, 5.5]).  The full operator is the sum of two engine applies, the
reciprocal part (mesh, order, xi) and the real-space and self parts (xi, rc),
so each is run once and every combination is formed from their rows.
Writes, per combination: system, N, phi, L, K, p, xi, rc and the relative
error of the 64 rows.

    python tests/golden/make_mesh_runs.py OUT.npz
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from spme_amd import engine as E  # noqa: E402
from spme_amd import params as PR  # noqa: E402
from spme_amd import workload as W  # noqa: E402

H_RANGE = (0.45, 1.35)
XIH_RANGE = (0.10, 0.80)
NXI, T_GRID = 14, np.linspace(1.5, 5.5, 13)


def meshes(L):
    ks = [k for k in PR.SMOOTH_K if H_RANGE[0] <= L / k <= H_RANGE[1]]
    if len(ks) > 3:
        ks = [ks[0], ks[len(ks) // 2], ks[-1]]
    return ks


def main(path):
    import torch
    g = np.load(os.path.join(HERE, "mesh_direct.npz"))
    rec = {k: [] for k in ("sys", "n", "phi", "L", "K", "p", "xi", "rc", "err")}
    t_all = time.time()
    for s in range(len(g["seed"])):
        seed, n, phi, L = int(g["seed"][s]), int(g["n"][s]), float(g["phi"][s]), float(g["L"][s])
        rows, direct = g["rows"][s], g["direct"][s]
        idx = torch.tensor((3 * rows[:, None] + np.arange(3)[None, :]).reshape(-1), device="cuda")
        pos = W.positions(seed, n, L)
        f = torch.tensor(W.forces(seed, n, 1), device="cuda")
        ks = meshes(L)
        hs = [L / k for k in ks]
        xis = np.geomspace(XIH_RANGE[0] / max(hs), XIH_RANGE[1] / min(hs), NXI)

        def rows_of(xi, K, p, rc, parts):
            eng = E.Engine(xi, K, p, n, L, rc=rc, precision=E.SPME_FP64, max_rhs=1)
            eng.set_positions(pos)
            u = eng.apply(f, parts=parts)
            out = u[0, idx].cpu().numpy()
            eng.close()
            return out

        real = {}
        for i, xi in enumerate(xis):
            for t in T_GRID:
                rc = float(t / xi)
                if rc <= L / 3.05:
                    real[(i, rc)] = rows_of(xi, 32, 4, rc, E.SPME_REAL | E.SPME_SELF)
        dn = np.linalg.norm(direct)
        for K in ks:
            for p in (4, 6, 8):
                if 4 * p >= K:
                    continue
                for i, xi in enumerate(xis):
                    if not XIH_RANGE[0] <= xi * L / K <= XIH_RANGE[1]:
                        continue
                    ur = rows_of(xi, K, p, 0.0, E.SPME_RECIP)
                    for (j, rc), uu in real.items():
                        if j != i:
                            continue
                        err = float(np.linalg.norm(ur + uu - direct) / dn)
                        for k, v in zip(rec, (s, n, phi, L, K, p, xi, rc, err)):
                            rec[k].append(v)
        print(f"system {s}: N={n} phi={phi:.3f} K={ks}, {len(rec['sys'])} combinations, "
              f"{time.time() - t_all:.0f}s", flush=True)
    np.savez_compressed(path, **{k: np.array(v) for k, v in rec.items()})
    print("wrote", path, flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
