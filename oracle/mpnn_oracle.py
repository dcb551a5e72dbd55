"""ORACLE — TEST INFRASTRUCTURE ONLY.

FP64 restatement of the learned initialiser's network, features and loss (no counterpart
in the reference): forward, the exact gradient of the loss (accumulated on both ends of
every ordered pair, so it does not rely on the pair set being symmetric), a BF16 rounding
model for the weights (round to nearest even), and the feature generator. Pinned by
central finite differences of its own forward (tests/test_mpnn_oracle.py) and by a
known-answer forward.
"""
import numpy as np

IN, HID = 16, 64
OFF = {}
_o = 0
for name, size in (("W1s", IN * HID), ("W1n", IN * HID), ("b1", HID), ("W2s", HID * HID), ("W2n", HID * HID),
                   ("b2", HID), ("Wo", HID * 2), ("bo", 2)):
    OFF[name] = (_o, size)
    _o += size
NPARAMS = _o
SHAPES = {"W1s": (IN, HID), "W1n": (IN, HID), "b1": (HID,), "W2s": (HID, HID), "W2n": (HID, HID), "b2": (HID,),
          "Wo": (HID, 2), "bo": (2,)}


def bf16_round(x):
    """FP32 -> BF16 (round to nearest even) -> float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    return {k: np.asarray(p[o:o + s], np.float64).reshape(SHAPES[k]) for k, (o, s) in OFF.items()}


def pack(d):
    out = np.zeros(NPARAMS)
    for k, (o, s) in OFF.items():
        out[o:o + s] = np.asarray(d[k]).ravel()
    return out


def mean_agg(rowptr, col, H):
    deg = np.diff(rowptr)
    rows = np.repeat(np.arange(len(deg)), deg)
    out = np.zeros((len(deg),) + H.shape[1:])
    np.add.at(out, rows, H[col])
    nz = deg > 0
    out[nz] /= deg[nz][:, None]
    return out


def mean_agg_T(rowptr, col, G):
    deg = np.diff(rowptr).astype(np.float64)
    rows = np.repeat(np.arange(len(deg)), np.diff(rowptr))
    out = np.zeros_like(G)
    # out[j] = sum_{i: j in N(i)} G[i]/deg(i)
    np.add.at(out, col, G[rows] / deg[rows][:, None])
    return out


def forward(p, rowptr, col, X, round_bf16=True):
    w = unpack(bf16_round(p) if round_bf16 else np.asarray(p, np.float64))
    X = np.asarray(X, np.float64)
    A0 = mean_agg(rowptr, col, X)
    Z1 = X @ w["W1s"] + A0 @ w["W1n"] + w["b1"]
    H1 = np.maximum(Z1, 0)
    A1 = mean_agg(rowptr, col, H1)
    Z2 = H1 @ w["W2s"] + A1 @ w["W2n"] + w["b2"]
    H2 = np.maximum(Z2, 0)
    Y = H2 @ w["Wo"] + w["bo"]
    return Y, {"w": w, "X": X, "A0": A0, "H1": H1, "A1": A1, "H2": H2}


def pairs(n, rowptr, col, offsets):
    deg = np.diff(rowptr)
    I = [np.repeat(np.arange(n), deg)]
    J = [np.asarray(col, np.int64)]
    for o in offsets:
        I += [np.arange(n), np.arange(n)]
        J += [(np.arange(n) + o) % n, (np.arange(n) - o) % n]
    return np.concatenate(I), np.concatenate(J)


def loss_grad(p, rowptr, col, X, T, offsets, round_bf16=True):
    """Loss (1/|P|) sum (|y_i-y_j| - |t_i-t_j|)^2 and its gradient w.r.t. p."""
    n = len(X)
    Y, c = forward(p, rowptr, col, X, round_bf16)
    I, J = pairs(n, rowptr, col, offsets)
    P = len(I)
    dv = Y[I] - Y[J]
    d = np.sqrt((dv ** 2).sum(1)) + 1e-12
    t = np.sqrt(((np.asarray(T, np.float64)[I] - np.asarray(T, np.float64)[J]) ** 2).sum(1))
    r = d - t
    L = (r ** 2).sum() / P
    g = (2.0 / P) * (r / d)[:, None] * dv
    dY = np.zeros_like(Y)
    np.add.at(dY, I, g)
    np.add.at(dY, J, -g)
    w = c["w"]
    grads = {"Wo": c["H2"].T @ dY, "bo": dY.sum(0)}
    dZ2 = (dY @ w["Wo"].T) * (c["H2"] > 0)
    grads["W2s"] = c["H1"].T @ dZ2
    grads["W2n"] = c["A1"].T @ dZ2
    grads["b2"] = dZ2.sum(0)
    dH1 = dZ2 @ w["W2s"].T + mean_agg_T(rowptr, col, dZ2 @ w["W2n"].T)
    dZ1 = dH1 * (c["H1"] > 0)
    grads["W1s"] = c["X"].T @ dZ1
    grads["W1n"] = c["A0"].T @ dZ1
    grads["b1"] = dZ1.sum(0)
    return L, pack(grads)


def mix64(z):
    z = np.asarray(z, np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def features(rowptr, col, seed, steps):
    """Column 0 log(deg+1); 1..8 seeded normal fields smoothed `steps` times by
    z <- (z + mean_agg(z))/2; all 9 standardised; 10..15 zero."""
    n = len(rowptr) - 1
    with np.errstate(over="ignore"):
        t = np.arange(n * 8, dtype=np.uint64)
        r = mix64(np.uint64(seed) ^ mix64(t))
    u1 = ((r >> np.uint64(40)).astype(np.float64) + 0.5) / 16777216.0
    u2 = ((r >> np.uint64(16)) & np.uint64(0xFFFFFF)).astype(np.float64) / 16777216.0
    z = (np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u2)).reshape(n, 8)
    for _ in range(steps):
        z = 0.5 * (z + mean_agg(rowptr, col, z))
    X = np.zeros((n, IN))
    X[:, 0] = np.log(np.diff(rowptr) + 1.0)
    X[:, 1:9] = z
    mu = X.mean(0)
    var = (X ** 2).mean(0) - mu ** 2
    sd = np.where((np.arange(IN) <= 8) & (var > 1e-12), 1.0 / np.sqrt(np.maximum(var, 1e-300)), 0.0)
    return (X - mu) * sd
