"""Pins the learned initialiser's FP64 oracle (CPU): its gradient against central finite
differences of its own forward, its forward on hand-set weights, and its BF16 model."""
import numpy as np
import pytest

from gpugraphlayout_amd import graphs
from oracle import mpnn_oracle as mo


def small_case(seed=0, n=40):
    s, t = graphs.barabasi_albert(n, 2, seed)
    g = graphs.edges_to_graph(s, t)
    rowptr, col = g.csr()
    X = mo.features(rowptr, col, seed=7, steps=3)
    rng = np.random.default_rng(seed)
    T = rng.normal(size=(g.n, 2))
    p = rng.normal(scale=0.3, size=mo.NPARAMS)
    return g, rowptr, col, X, T, p


def test_gradient_matches_finite_differences():
    g, rowptr, col, X, T, p = small_case()
    offs = [3, 11]
    L, grad = mo.loss_grad(p, rowptr, col, X, T, offs, round_bf16=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(np.arange(o, o + s), 4, replace=False) for o, s in mo.OFF.values()
                          if s > 2] + [np.arange(mo.OFF["bo"][0], mo.OFF["bo"][0] + 2)])
    # only weights that touch live inputs (columns 9..15 of X are zero)
    for k in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mo.loss_grad(pp, rowptr, col, X, T, offs, False)[0] - mo.loss_grad(pm, rowptr, col, X, T, offs, False)[0]) / (2 * h)
        assert grad[k] == pytest.approx(fd, rel=1e-5, abs=1e-9), k


def test_forward_known_answer():
    # hidden unit 0 = relu(x0); unit 1 = relu(mean of neighbours' x0); y = (h0 + h1, -h1)
    n = 4
    rowptr = np.array([0, 1, 3, 5, 6])
    col = np.array([1, 0, 2, 1, 3, 2])        # path 0-1-2-3
    X = np.zeros((n, mo.IN))
    X[:, 0] = [1.0, -2.0, 3.0, 0.5]
    w = {k: np.zeros(s) for k, s in mo.SHAPES.items()}
    w["W1s"][0, 0] = 1.0
    w["W1n"][0, 1] = 1.0
    w["W2s"][0, 0] = 1.0
    w["W2s"][1, 1] = 1.0
    w["Wo"][0, 0] = 1.0
    w["Wo"][1, 0] = 1.0
    w["Wo"][1, 1] = -1.0
    Y, _ = mo.forward(mo.pack(w), rowptr, col, X)
    h0 = np.maximum(X[:, 0], 0)
    nb = np.array([-2.0, (1 + 3) / 2, (-2 + 0.5) / 2, 3.0])
    h1 = np.maximum(nb, 0)
    assert np.allclose(Y, np.stack([h0 + h1, -h1], 1))


def test_bf16_rounding_model():
    x = np.array([1.0, 1.00390625, 1.005859375, 3.14159265, -2.5e-3], np.float32)
    r = mo.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0            # exactly halfway: ties to even
    assert r[2] == 1.0078125                       # above halfway rounds up
    assert abs(r[3] - 3.140625) < 1e-12
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)


def test_loss_invariances():
    g, rowptr, col, X, T, p = small_case(2)
    L1, _ = mo.loss_grad(p, rowptr, col, X, T, [5])
    c, s_ = np.cos(0.7), np.sin(0.7)
    T2 = T @ np.array([[c, -s_], [s_, c]]) * np.array([1, -1]) + 3.0   # rotate, reflect, translate
    L2, _ = mo.loss_grad(p, rowptr, col, X, T2, [5])
    assert L1 == pytest.approx(L2, rel=1e-10)


def test_features_standardised():
    g, rowptr, col, X, T, p = small_case(3, n=300)
    assert np.allclose(X[:, :9].mean(0), 0, atol=1e-10)
    assert np.allclose(X[:, :9].std(0), 1, atol=1e-8)
    assert np.all(X[:, 9:] == 0)
