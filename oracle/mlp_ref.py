"""CPU restatement (float64 numpy) of the rebuild-interval network of
rockable_amd/csrc/rk_mlp.hip: 16 -> 64 (tanh) -> 64 (tanh) -> 1, loss = mean
of (y - t)^2.  Test infrastructure: it checks the device's forward pass and
gradients; its own gradient is checked against finite differences."""
import numpy as np

NIN, NH = 16, 64
SHAPES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH,)), ("b3", (1,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = p[o:o + n].reshape(shp)
        o += n
    return out


def forward(p, X):
    w = unpack(p)
    X = np.asarray(X, dtype=np.float64).reshape(-1, NIN)
    h1 = np.tanh(X @ w["W1"].T + w["b1"])
    h2 = np.tanh(h1 @ w["W2"].T + w["b2"])
    return h2 @ w["W3"] + w["b3"][0], (X, h1, h2)


def loss_grad(p, X, t):
    w = unpack(p)
    y, (X, h1, h2) = forward(p, X)
    t = np.asarray(t, dtype=np.float64)
    n = len(t)
    r = y - t
    loss = float(np.mean(r * r))
    e = 2.0 * r / n                                  # dL/dy
    d2 = (e[:, None] * w["W3"][None, :]) * (1.0 - h2 * h2)
    d1 = (d2 @ w["W2"]) * (1.0 - h1 * h1)
    g = [d1.T @ X, d1.sum(0), d2.T @ h1, d2.sum(0), e @ h2, np.array([e.sum()])]
    return loss, np.concatenate([x.ravel() for x in g])
