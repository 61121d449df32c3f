"""CPU FP64 restatement of the parameter-selection network and its gradient
(spme_amd/csrc/spme_mlp.hip): forward, loss, backward.

TEST INFRASTRUCTURE ONLY (the checker of the GPU kernels).  Pinned in
tests/test_mlp.py by central finite differences of the loss.
"""
import numpy as np

NI, NH, NO = 4, 64, 4
SIZES = [("W1", (NH, NI)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
         ("W3", (NO, NH)), ("b3", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(w):
    out, o = {}, 0
    for name, shp in SIZES:
        n = int(np.prod(shp))
        out[name] = np.asarray(w[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def forward(w, x):
    P = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.tanh(x @ P["W1"].T + P["b1"])
    h2 = np.tanh(h1 @ P["W2"].T + P["b2"])
    y = h2 @ P["W3"].T + P["b3"]
    return y, h1, h2


def loss(w, x, t, mask):
    y, _, _ = forward(w, x)
    return float(np.sum(np.asarray(mask) * (y - t) ** 2) / x.shape[0])


def grad(w, x, t, mask):
    P = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    B = x.shape[0]
    y, h1, h2 = forward(w, x)
    dy = 2.0 * (y - t) * np.asarray(mask) / B
    d2 = (dy @ P["W3"]) * (1 - h2 ** 2)
    d1 = (d2 @ P["W2"]) * (1 - h1 ** 2)
    g = {"W3": dy.T @ h2, "b3": dy.sum(0), "W2": d2.T @ h1, "b2": d2.sum(0),
         "W1": d1.T @ x, "b1": d1.sum(0)}
    return np.concatenate([g[n].ravel() for n, _ in SIZES])
