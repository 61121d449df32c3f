"""Learned parameter selection: a 4 -> 64 -> 64 -> 4 MLP trained and run by
the engine's own HIP kernels (spme_amd/csrc/spme_mlp.hip).

Inputs  (log N, phi, log10 tol, L/a), standardised.
Outputs (log xi, log2 K, p, log rc), standardised; K is snapped to the
        smooth set {32, 48, 64, 96, 128, 192, 256}, p to {4, 6, 8}.

Labels come from measured SPME runs (tests/golden/make_mlp_data.py): for
each system and tolerance, the fastest evaluated parameter set whose
measured relative error against the FP64 direct Ewald sum is within the
tolerance.  The analytic estimate (spme_amd/params.py) is the fallback when
no trained weights are present or when the prediction is invalid.
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
NI, NH, NO = 4, 64, 4


def features(n, phi, tol):
    L = PR.box_length(n, phi)
    return np.array([math.log(n), phi, math.log10(tol), L], dtype=np.float64)


def targets(xi, K, p, rc):
    return np.array([math.log(xi), math.log2(K), float(p), math.log(rc)], dtype=np.float64)


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
    def in_domain(self, n, phi, tol):
        """True inside the (N, phi, tol) box the network was trained on; the
        network is not trusted to extrapolate outside it."""
        d = getattr(self, "meta", {}).get("domain")
        if not d:
            return True
        return (d["n"][0] <= n <= d["n"][1] and d["phi"][0] <= phi <= d["phi"][1]
                and d["tol"][0] <= tol <= d["tol"][1])

    def predict(self, n, phi, tol):
        xs = (features(n, phi, tol) - self.x_mean) / self.x_std
        ys = self.forward_raw(xs[None, :])[0] * self.y_std + self.y_mean
        return ys

    def select(self, n, phi, tol, nrhs=1, k_max=256):
        """(xi, K, p, rc) for the target relative error tol; falls back to the
        analytic estimate if the prediction is unusable."""
        L = PR.box_length(n, phi)
        lxi, lk, pp, lrc = self.predict(n, phi, tol)
        xi, rc = math.exp(lxi), math.exp(lrc)
        Kc = [k for k in PR.SMOOTH_K if k <= k_max]
        K = min(Kc, key=lambda k: abs(math.log2(k) - lk))
        p = min(PR.ORDERS, key=lambda q: abs(q - pp))
        while 4 * p >= K and p > 4:
            p -= 2
        ok = (0.05 < xi < 3.0) and (0.5 < rc <= L / 3.0) and 4 * p < K
        if not ok or not self.in_domain(n, phi, tol):
            a = PR.analytic_select(n, phi, tol, nrhs=nrhs, k_max=k_max)
            return a.xi, a.K, a.p, a.rc, "analytic"
        return xi, K, p, rc, "learned"

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
    within tol / margin.  runs: dict of equal-length arrays with keys
    sys, n, phi, xi, K, p, rc, err, time.  Returns (X, T, info)."""
    X, T, info = [], [], []
    sysids = np.unique(runs["sys"])
    for s in sysids:
        sel = np.where(runs["sys"] == s)[0]
        for tol in tols:
            ok = sel[runs["err"][sel] <= tol / margin]
            if ok.size == 0:
                continue
            b = ok[np.argmin(runs["time"][ok])]
            X.append(features(int(runs["n"][b]), float(runs["phi"][b]), tol))
            T.append(targets(runs["xi"][b], runs["K"][b], runs["p"][b], runs["rc"][b]))
            info.append((int(s), tol, int(b)))
    return np.array(X), np.array(T), info


def training_domain(runs, tols=TOLS):
    """The (N, phi, tol) box covered by the training runs."""
    return {"n": [int(runs["n"].min()), int(runs["n"].max())],
            "phi": [float(runs["phi"].min()), float(runs["phi"].max())],
            "tol": [float(min(tols)), float(max(tols))]}


def fit(runs, epochs=4000, lr=3e-3, seed=0, log_every=500, margin=1.5):
    """Train a network on measured runs; returns (net, history).  Labels keep
    a margin below the tolerance so that predictions, which interpolate
    between neighbouring systems, still land inside it."""
    X, T, _ = labels_from_runs(runs, margin=margin)
    xm, xsd = X.mean(0), X.std(0) + 1e-12
    ym, ysd = T.mean(0), T.std(0) + 1e-12
    net = Net(init_params(seed), xm, xsd, ym, ysd)
    hist = net.train((X - xm) / xsd, (T - ym) / ysd, epochs=epochs, lr=lr,
                     log_every=log_every)
    net.meta = {"samples": int(len(X)), "epochs": epochs, "lr": lr, "seed": seed,
                "margin": margin, "domain": training_domain(runs),
                "final_loss": hist[-1][1] if hist else None}
    return net, hist


def select(n, phi, tol, nrhs=1, k_max=256):
    """Learned selection if trained weights are present, else analytic."""
    if os.path.exists(WEIGHTS):
        return Net.load().select(n, phi, tol, nrhs=nrhs, k_max=k_max)
    a = PR.analytic_select(n, phi, tol, nrhs=nrhs, k_max=k_max)
    return a.xi, a.K, a.p, a.rc, "analytic"


def load_runs(*paths):
    """Concatenate run files written by the training-data generator."""
    parts = [np.load(p) for p in paths]
    return {k: np.concatenate([z[k] for z in parts]) for k in parts[0].files}


if __name__ == "__main__":
    # python -m spme_amd.mlp OUT.npz RUNS.npz [RUNS.npz ...]
    import sys
    runs = load_runs(*sys.argv[2:])
    net, hist = fit(runs)
    for ep, loss in hist:
        print(f"epoch {ep} loss {loss:.4e}", flush=True)
    net.meta["runs"] = int(len(runs["err"]))
    net.save(sys.argv[1], meta=net.meta)
    print("saved", sys.argv[1], net.meta, flush=True)
