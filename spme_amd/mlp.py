"""Learned parameter selection: 4 -> 64 -> 64 -> 4 MLPs trained and run by
the engine's own HIP kernels (spme_amd/csrc/spme_mlp.hip).

The tolerance enters as an absolute tolerance, tol_abs = tol * rms(u)/rms(f)
with rms(u)/rms(f) = sqrt(1 + 0.66 phi L) (spme_amd/params.py): the relative
error of a fixed parameter set falls as the box grows (the long-range 1/r
field grows with it), while the absolute error per unit force is intensive.

Full selector (mesh, order, xi and rc free):
  inputs  (log N, phi, log10 tol_abs, L/a), standardised;
  outputs (log xi, log h, p, log rc), standardised, with h = L/K the mesh
          spacing: K is the smooth size in {32, 48, 64, 96, 128, 192, 256}
          nearest to L/h in log (the labels are such sizes), p is snapped to
          {4, 6, 8}.
  Labels come from measured SPME runs (tests/golden/make_mlp_data.py): for
  each system and tolerance, the fastest evaluated parameter set whose
  measured relative error against the FP64 direct Ewald sum is within the
  tolerance.  At fixed (xi, h, p, rc) SPME's absolute error and its
  per-particle cost are intensive, so the outputs carry over to larger
  systems: inputs outside the training box are clamped to its edge (N
  clamped and L recomputed from it and phi, so the inputs stay a physical
  system; tol_abs from the true box, then clamped), and when the mesh the
  spacing asks for exceeds the largest one allowed, xi h and xi rc are kept.

Fixed-mesh selector (K and p given, xi and rc free):
  inputs  (phi, log10 tol_abs, h, p), standardised; outputs (log xi, log rc)
          (the last two outputs are unused and masked out of the loss).
  Labels come from runs over a grid of xi and rc at fixed meshes
  (tests/golden/make_mesh_runs.py): the xi whose measured error falls within
  the tolerance at the smallest rc (at a fixed mesh the cost grows with rc
  only).

The analytic estimate (spme_amd/params.py) is the fallback when no trained
weights are present or when a prediction is unusable.
"""
from __future__ import annotations

import ctypes
import json
import math
import os

import numpy as np

from . import params as PR
from .lib import check, load

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "mlp_weights.npz")
MESH_WEIGHTS = os.path.join(HERE, "data", "mlp_mesh_weights.npz")
NI, NH, NO = 4, 64, 4


def tol_abs(tol, phi, L):
    """Absolute tolerance per unit rms force of a relative tolerance."""
    return tol * PR.rms_u(phi, L)


def features(n, phi, tol, L=None):
    """(log N, phi, log10 tol_abs, L); tol_abs from the box L (default: the
    box of N and phi)."""
    Lb = PR.box_length(n, phi)
    return np.array([math.log(n), phi, math.log10(tol_abs(tol, phi, Lb if L is None else L)),
                     Lb], dtype=np.float64)


def targets(xi, K, p, rc, L):
    return np.array([math.log(xi), math.log(L / K), float(p), math.log(rc)], dtype=np.float64)


def mesh_features(phi, tol, h, p, L):
    return np.array([phi, math.log10(tol_abs(tol, phi, L)), h, float(p)], dtype=np.float64)


def mesh_targets(xi, rc):
    return np.array([math.log(xi), math.log(rc), 0.0, 0.0], dtype=np.float64)


MESH_MASK = (1.0, 1.0, 0.0, 0.0)


def _clamp(v, lo, hi):
    return min(max(v, lo), hi)


def mesh_for(L, h, k_max=256):
    """The smooth K <= k_max nearest to L/h in log."""
    ks = [k for k in PR.SMOOTH_K if k <= k_max]
    return min(ks, key=lambda k: abs(math.log(k * h / L)))


def nparam() -> int:
    return load().spme_mlp_nparam()


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases (FP32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_out, fan_in in [(NH, NI), (NH, NH), (NO, NH)]:
        lim = math.sqrt(6.0 / (fan_in + fan_out))
        parts.append(rng.uniform(-lim, lim, fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    w = np.concatenate(parts).astype(np.float32)
    assert w.size == nparam()
    return w


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Net:
    """Weights on the GPU plus the standardisation of inputs and outputs."""

    def __init__(self, w, x_mean, x_std, y_mean, y_std):
        import torch
        self.w = torch.tensor(np.asarray(w, np.float32), device="cuda")
        self.x_mean, self.x_std = np.asarray(x_mean), np.asarray(x_std)
        self.y_mean, self.y_std = np.asarray(y_mean), np.asarray(y_std)

    # -- kernels ---------------------------------------------------------
    def forward_raw(self, xs):
        """xs: standardised inputs [B, 4] (numpy) -> standardised outputs."""
        import torch
        x = torch.tensor(np.asarray(xs, np.float32), device="cuda").contiguous()
        y = torch.empty((x.shape[0], NO), dtype=torch.float32, device="cuda")
        check(load().spme_mlp_forward(_ptr(self.w), _ptr(x), int(x.shape[0]), _ptr(y), None,
                                      None, _stream()), "mlp forward")
        return y.cpu().numpy()

    def train(self, X, T, mask=(1, 1, 1, 1), epochs=3000, lr=3e-3, log_every=0):
        """Full-batch Adam on standardised (X, T); returns the loss history."""
        import torch
        L = load()
        x = torch.tensor(np.asarray(X, np.float32), device="cuda").contiguous()
        t = torch.tensor(np.asarray(T, np.float32), device="cuda").contiguous()
        m = torch.tensor(np.asarray(mask, np.float32), device="cuda")
        B = x.shape[0]
        y = torch.empty((B, NO), dtype=torch.float32, device="cuda")
        h1 = torch.empty((B, NH), dtype=torch.float32, device="cuda")
        h2 = torch.empty_like(h1)
        work = torch.empty(B * (NO + 2 * NH), dtype=torch.float32, device="cuda")
        g = torch.zeros_like(self.w)
        m1 = torch.zeros_like(self.w)
        m2 = torch.zeros_like(self.w)
        st = _stream()
        hist = []
        for ep in range(1, epochs + 1):
            check(L.spme_mlp_forward(_ptr(self.w), _ptr(x), B, _ptr(y), _ptr(h1), _ptr(h2),
                                     st), "mlp forward")
            check(L.spme_mlp_backward(_ptr(self.w), _ptr(x), _ptr(y), _ptr(t), _ptr(m), B,
                                      _ptr(h1), _ptr(h2), _ptr(work), _ptr(g), st),
                  "mlp backward")
            check(L.spme_mlp_adam(_ptr(self.w), _ptr(g), _ptr(m1), _ptr(m2), float(lr), 0.9,
                                  0.999, 1e-8, ep, st), "mlp adam")
            if log_every and (ep % log_every == 0 or ep == epochs):
                loss = float(((y - t) ** 2 * m).sum().item() / B)
                hist.append((ep, loss))
        torch.cuda.synchronize()
        return hist

    # -- selection -------------------------------------------------------
    def _domain(self):
        return getattr(self, "meta", {}).get("domain") or {}

    def predict_x(self, x):
        xs = (np.asarray(x, np.float64) - self.x_mean) / self.x_std
        return self.forward_raw(xs[None, :])[0] * self.y_std + self.y_mean

    def predict(self, n, phi, tol):
        """Raw outputs (log xi, log h, p, log rc) for a system, its inputs
        clamped to the training box (N clamped and L recomputed from it;
        tol_abs from the true box, then clamped)."""
        x = features(n, phi, tol)
        d = self._domain()
        if d:
            nc, phic = _clamp(n, *d["n"]), _clamp(phi, *d["phi"])
            x = features(nc, phic, 1.0)
            x[2] = _clamp(math.log10(tol_abs(tol, phi, PR.box_length(n, phi))),
                          *d["log10_tol_abs"])
        return self.predict_x(x)

    def select(self, n, phi, tol, nrhs=1, k_max=256):
        """(xi, K, p, rc, how) for the target relative error tol (full
        selector); how is "learned", or "analytic" for the fallback."""
        L = PR.box_length(n, phi)
        lxi, lh, pp, lrc = self.predict(n, phi, tol)
        xi, h, rc = math.exp(lxi), math.exp(lh), math.exp(lrc)
        K = mesh_for(L, h, k_max)
        if K == max(k for k in PR.SMOOTH_K if k <= k_max) and L / K > h:
            # the mesh is capped: keep xi h (mesh error) and xi rc (real space)
            s = h / (L / K)
            xi, rc = xi * s, rc / s
        p = min(PR.ORDERS, key=lambda q: abs(q - pp))
        while 4 * p >= K and p > 4:
            p -= 2
        ok = (0.05 < xi < 3.0) and (0.5 < rc <= L / 3.0) and 4 * p < K
        if not ok:
            a = PR.analytic_select(n, phi, tol, nrhs=nrhs, k_max=k_max)
            return a.xi, a.K, a.p, a.rc, "analytic"
        return xi, K, p, rc, "learned"

    def select_mesh(self, n, phi, tol, K, p):
        """(xi, rc, how) at a given mesh and order (fixed-mesh selector)."""
        L = PR.box_length(n, phi)
        x = mesh_features(phi, tol, L / K, p, L)
        d = self._domain()
        if d:
            x = np.array([_clamp(x[0], *d["phi"]), _clamp(x[1], *d["log10_tol_abs"]),
                          _clamp(x[2], *d["h"]), x[3]])
        lxi, lrc = self.predict_x(x)[:2]
        xi, rc = math.exp(lxi), math.exp(lrc)
        if not ((0.05 < xi < 3.0) and (0.5 < rc <= L / 3.0)):
            a = PR.analytic_select(n, phi, tol, fixed_K=K, fixed_p=p)
            return a.xi, a.rc, "analytic"
        return xi, rc, "learned"

    # -- storage -----------------------------------------------------------
    def save(self, path=WEIGHTS, meta=None):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, w=self.w.cpu().numpy(), x_mean=self.x_mean, x_std=self.x_std,
                 y_mean=self.y_mean, y_std=self.y_std,
                 meta=json.dumps(meta or {}))

    @classmethod
    def load(cls, path=WEIGHTS):
        z = np.load(path, allow_pickle=False)
        net = cls(z["w"], z["x_mean"], z["x_std"], z["y_mean"], z["y_std"])
        net.meta = json.loads(str(z["meta"]))
        return net

    def set_meta(self, **kw):
        self.meta = dict(getattr(self, "meta", {}), **kw)


TOLS = (1e-3, 3e-4, 1e-4, 3e-5, 1e-5, 3e-6, 1e-6, 3e-7, 1e-7)


def labels_from_runs(runs, tols=TOLS, margin=1.0):
    """For every system and tolerance: the fastest run whose measured error is
    within tol / margin.  runs: dict of equal-length arrays with keys sys, n,
    phi, xi, K, p, rc, err, time.  Returns (X, T, info)."""
    X, T, info = [], [], []
    for s in np.unique(runs["sys"]):
        sel = np.where(runs["sys"] == s)[0]
        for tol in tols:
            ok = sel[runs["err"][sel] <= tol / margin]
            if ok.size == 0:
                continue
            b = ok[np.argmin(runs["time"][ok])]
            n, phi = int(runs["n"][b]), float(runs["phi"][b])
            X.append(features(n, phi, tol))
            T.append(targets(runs["xi"][b], runs["K"][b], runs["p"][b], runs["rc"][b],
                             PR.box_length(n, phi)))
            info.append((int(s), tol, int(b)))
    return np.array(X), np.array(T), info


def training_domain(X, runs):
    """The box covered by the training runs: N, phi and log10 tol_abs."""
    return {"n": [int(runs["n"].min()), int(runs["n"].max())],
            "phi": [float(runs["phi"].min()), float(runs["phi"].max())],
            "log10_tol_abs": [float(X[:, 2].min()), float(X[:, 2].max())]}


def mesh_labels(runs, tols=TOLS, margin=1.5):
    """Fixed-mesh labels.  For every (system, K, p), tolerance and xi of the
    grid: the smallest rc from which on the measured error stays within
    tol / margin, log-interpolated between the two grid cutoffs around the
    crossing; the label is the xi with the smallest such rc.  runs: keys
    sys, n, phi, L, K, p, xi, rc, err.  Returns (X, T, info)."""
    X, T, info = [], [], []
    key = np.stack([runs["sys"], runs["K"], runs["p"]], 1)
    for k in np.unique(key, axis=0):
        sel = np.where((key == k).all(1))[0]
        s, K, p = (int(v) for v in k)
        L, phi = float(runs["L"][sel[0]]), float(runs["phi"][sel[0]])
        xis = np.unique(runs["xi"][sel])
        for tol in tols:
            target = tol / margin
            best = None
            for xi in xis:
                q = sel[runs["xi"][sel] == xi]
                q = q[np.argsort(runs["rc"][q])]
                rc, err = runs["rc"][q], runs["err"][q]
                above = np.where(err > target)[0]
                i = 0 if above.size == 0 else above[-1] + 1
                if i >= rc.size:
                    continue
                r = rc[i]
                if i > 0:  # crossing between rc[i-1] (above) and rc[i]
                    la, lb = math.log(err[i - 1]), math.log(max(err[i], 1e-300))
                    t = (la - math.log(target)) / (la - lb) if la > lb else 1.0
                    r = rc[i - 1] + t * (rc[i] - rc[i - 1])
                if best is None or r < best[1]:
                    best = (float(xi), float(r))
            if best is None:
                continue
            X.append(mesh_features(phi, tol, L / K, p, L))
            T.append(mesh_targets(*best))
            info.append((s, K, p, tol))
    return np.array(X), np.array(T), info


def _train(X, T, mask, epochs, lr, seed, log_every):
    xm, xsd = X.mean(0), X.std(0) + 1e-12
    ym, ysd = T.mean(0), T.std(0) + 1e-12
    net = Net(init_params(seed), xm, xsd, ym, ysd)
    hist = net.train((X - xm) / xsd, (T - ym) / ysd, mask=mask, epochs=epochs, lr=lr,
                     log_every=log_every)
    return net, hist


def fit(runs, epochs=4000, lr=3e-3, seed=0, log_every=500, margin=1.5):
    """Train the full selector on measured runs; returns (net, history).
    Labels keep a margin below the tolerance so that predictions, which
    interpolate between neighbouring systems, still land inside it."""
    X, T, _ = labels_from_runs(runs, margin=margin)
    net, hist = _train(X, T, (1, 1, 1, 1), epochs, lr, seed, log_every)
    net.meta = {"kind": "full", "samples": int(len(X)), "epochs": epochs, "lr": lr,
                "seed": seed, "margin": margin, "domain": training_domain(X, runs),
                "final_loss": hist[-1][1] if hist else None}
    return net, hist


def fit_mesh(runs, epochs=4000, lr=3e-3, seed=0, log_every=500, margin=1.25):
    """Train the fixed-mesh selector on grid runs; returns (net, history).
    Its labels are read off dense (xi, rc) grids, at the tolerance itself
    rather than at the best of a few candidates, so a smaller margin than
    the full selector's holds them inside it."""
    X, T, _ = mesh_labels(runs, margin=margin)
    net, hist = _train(X, T, MESH_MASK, epochs, lr, seed, log_every)
    h = runs["L"] / runs["K"]
    net.meta = {"kind": "mesh", "samples": int(len(X)), "epochs": epochs, "lr": lr,
                "seed": seed, "margin": margin,
                "domain": {"phi": [float(runs["phi"].min()), float(runs["phi"].max())],
                           "log10_tol_abs": [float(X[:, 1].min()), float(X[:, 1].max())],
                           "h": [float(h.min()), float(h.max())]},
                "final_loss": hist[-1][1] if hist else None}
    return net, hist


def select(n, phi, tol, nrhs=1, k_max=256):
    """Learned selection if trained weights are present, else analytic."""
    if os.path.exists(WEIGHTS):
        return Net.load().select(n, phi, tol, nrhs=nrhs, k_max=k_max)
    a = PR.analytic_select(n, phi, tol, nrhs=nrhs, k_max=k_max)
    return a.xi, a.K, a.p, a.rc, "analytic"


def select_mesh(n, phi, tol, K, p):
    """(xi, rc, how) at a fixed mesh and order: the fixed-mesh network if its
    weights are present, else the analytic estimate."""
    if os.path.exists(MESH_WEIGHTS):
        return Net.load(MESH_WEIGHTS).select_mesh(n, phi, tol, K, p)
    a = PR.analytic_select(n, phi, tol, fixed_K=K, fixed_p=p)
    return a.xi, a.rc, "analytic"


def load_runs(*paths):
    """Concatenate run files written by the training-data generator."""
    parts = [np.load(p) for p in paths]
    return {k: np.concatenate([z[k] for z in parts]) for k in parts[0].files}


if __name__ == "__main__":
    # python -m spme_amd.mlp full|mesh OUT.npz RUNS.npz [RUNS.npz ...]
    import sys
    kind, out = sys.argv[1], sys.argv[2]
    runs = load_runs(*sys.argv[3:])
    net, hist = (fit if kind == "full" else fit_mesh)(runs)
    for ep, loss in hist:
        print(f"epoch {ep} loss {loss:.4e}", flush=True)
    net.meta["runs"] = int(len(runs["err"]))
    net.save(out, meta=net.meta)
    print("saved", out, net.meta, flush=True)
