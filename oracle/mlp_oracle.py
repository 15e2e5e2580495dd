"""CPU oracle of the learned wall model (numpy).

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its
C-ABI or the host package) imports this module.  It is used by tests/ as
the checker of the GPU forward/backward kernels.

The reference has no learned wall model, so there is no reference line to
restate; the oracle restates the network the library implements:

    x (4 standardised features) -> tanh(x W1 + b1) -> tanh(. W2 + b2)
      -> tanh(. W3 + b3) -> . W4 + b4 = y        (weights stored [in][out])

loss = sum_s (y_s - Y_s)^2, gradients of loss / B (mean squared error), and
the features / target standardisation

    f = [ln u, ln h, ln nu, asinh(dpdx / 1e-3)],  X = (f - mean) / std,
    Y = (ln tau - mean_tau) / std_tau,            tau = exp(y std_tau + mean_tau).

With ``bf16=True`` the operands of the two 64x64 products (activations,
weights and back-propagated errors) are rounded to bfloat16 (round to
nearest even) at the points where the device rounds them; sums are exact
products accumulated in float64 (the device accumulates in float32).
With ``bf16=False`` everything is float64; that mode is pinned against
central finite differences of the loss.
"""
import numpy as np

H = 64
LAYOUT = (("W1", (4, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
          ("W3", (H, H)), ("b3", (H,)), ("W4", (H,)), ("b4", (1,)))
NPARAMS = sum(int(np.prod(s)) for _, s in LAYOUT)


def unpack(theta):
    out, o = {}, 0
    for name, shp in LAYOUT:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[n], dtype=np.float64).ravel() for n, _ in LAYOUT])


def bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even); returns float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def raw_features(feat):
    """[ln u, ln h, ln nu, asinh(dpdx/1e-3)] of raw rows (u, h, nu, dpdx)."""
    feat = np.asarray(feat, dtype=np.float64)
    return np.stack([np.log(feat[:, 0]), np.log(feat[:, 1]), np.log(feat[:, 2]),
                     np.arcsinh(feat[:, 3] / 1e-3)], axis=1)


def fit_norm(feat, tau):
    """Means / standard deviations (population) and the feature box."""
    r = raw_features(feat)
    lt = np.log(np.asarray(tau, dtype=np.float64))
    nrm = np.zeros(16, dtype=np.float32)
    nrm[0:4] = r.mean(axis=0)
    sd = r.std(axis=0)
    nrm[4:8] = np.where(sd > 1e-12, sd, 1.0)
    nrm[8] = lt.mean()
    nrm[9] = lt.std()
    X = standardise(feat, nrm)
    hull = np.concatenate([X.min(axis=0), X.max(axis=0)]).astype(np.float32)
    return nrm, hull


def standardise(feat, nrm):
    r = raw_features(feat)
    return ((r - nrm[0:4].astype(np.float64)) / nrm[4:8].astype(np.float64)).astype(np.float32)


def target(tau, nrm):
    return ((np.log(np.asarray(tau, dtype=np.float64)) - float(nrm[8])) / float(nrm[9])).astype(np.float32)


def forward(theta, X, bf16_ops=True):
    """Network output y (float64) and the activations."""
    p = unpack(theta)
    rd = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    X = np.asarray(X, dtype=np.float64)
    a1 = np.tanh(X @ p["W1"] + p["b1"])
    a2 = np.tanh(rd(a1) @ rd(p["W2"]) + p["b2"])
    a3 = np.tanh(rd(a2) @ rd(p["W3"]) + p["b3"])
    y = a3 @ p["W4"] + p["b4"][0]
    return y, (X, a1, a2, a3)


def loss_and_grad(theta, X, Y, bf16_ops=True):
    """(sum of squared errors, gradient of the mean squared error)."""
    p = unpack(theta)
    rd = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    y, (X, a1, a2, a3) = forward(theta, X, bf16_ops)
    B = X.shape[0]
    e = y - np.asarray(Y, dtype=np.float64)
    dy = 2.0 * e / B
    g = {}
    g["W4"] = a3.T @ dy
    g["b4"] = np.array([dy.sum()])
    dz3 = dy[:, None] * p["W4"][None, :] * (1.0 - a3 * a3)
    g["b3"] = dz3.sum(axis=0)
    g["W3"] = rd(a2).T @ rd(dz3)
    dz2 = (rd(dz3) @ rd(p["W3"]).T) * (1.0 - a2 * a2)
    g["b2"] = dz2.sum(axis=0)
    g["W2"] = rd(a1).T @ rd(dz2)
    dz1 = (rd(dz2) @ rd(p["W2"]).T) * (1.0 - a1 * a1)
    g["b1"] = dz1.sum(axis=0)
    g["W1"] = X.T @ dz1
    return float((e * e).sum()), pack(g)


def predict_tau(theta, feat, nrm, bf16_ops=True):
    y, _ = forward(theta, standardise(feat, nrm), bf16_ops)
    return np.exp(y * float(nrm[9]) + float(nrm[8]))


def inside_hull(feat, nrm, hull):
    X = standardise(feat, nrm)
    ok = np.all((X >= hull[0:4]) & (X <= hull[4:8]), axis=1)
    return ok & (np.asarray(feat)[:, 0] > 0)
