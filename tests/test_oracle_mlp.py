"""Pins the CPU restatement of the order-selection MLP (oracle/mlp.py).

The model has no reference counterpart; its oracle is pinned by central
finite differences of its own loss in float64 (without the bfloat16
rounding, which makes the loss piecewise constant in the weights), by the
bfloat16 rounding rule on hand-picked values, and by closed forms of the
cross entropy.
"""
import numpy as np

import mlp


def _weights(seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(mlp.N_WEIGHTS, np.float32)
    w[: mlp.N_PARAMS] = rng.normal(0, 0.3, mlp.N_PARAMS)
    w[mlp.O_MU:mlp.O_SD] = rng.normal(0, 1, mlp.N_IN)
    w[mlp.O_SD:] = rng.uniform(0.5, 2.0, mlp.N_IN)
    return w


def test_layout_sizes():
    assert mlp.N_PARAMS == 5963 and mlp.N_WEIGHTS == 5995


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 2.0 ** -7, 1.0 + 3 * 2.0 ** -8, -2.5, 3.0e-3], np.float32)
    r = mlp.bf16_round(v)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2.0 ** -7)
    assert r[3] == np.float32(1.0 + 2.0 ** -6) and r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2.0 ** -8
    # 8 significant bits: the low 16 bits are zero
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    w = _weights(1).astype(np.float64)
    feat = rng.normal(0, 1.5, (40, mlp.N_IN))
    lab = rng.integers(0, mlp.N_OUT, 40)
    _, g = mlp.loss_grad(w, feat, lab, bf16=False)

    def L(wv):
        P = {
            "W1": wv[mlp.O_W1:mlp.O_B1].reshape(mlp.N_IN, mlp.N_H), "b1": wv[mlp.O_B1:mlp.O_W2],
            "W2": wv[mlp.O_W2:mlp.O_B2].reshape(mlp.N_H, mlp.N_H), "b2": wv[mlp.O_B2:mlp.O_W3],
            "W3": wv[mlp.O_W3:mlp.O_B3].reshape(mlp.N_H, mlp.N_OUT), "b3": wv[mlp.O_B3:mlp.O_MU],
            "mu": wv[mlp.O_MU:mlp.O_SD], "sd": wv[mlp.O_SD:],
        }
        z = mlp.forward(P, feat)[3]
        p = mlp.softmax(z)
        return -np.log(p[np.arange(len(lab)), lab]).mean()

    # one index in every parameter block, plus random ones
    idx = [mlp.O_W1 + 5, mlp.O_B1 + 3, mlp.O_W2 + 700, mlp.O_B2 + 10, mlp.O_W3 + 30, mlp.O_B3 + 2]
    idx += list(rng.integers(0, mlp.N_PARAMS, 30))
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (L(wp) - L(wm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_cross_entropy_closed_forms():
    # zero weights: uniform probabilities, loss ln 11, bias gradient (1/11 - onehot) averaged
    w = np.zeros(mlp.N_WEIGHTS, np.float32)
    w[mlp.O_SD:] = 1.0
    feat = np.random.default_rng(2).normal(size=(7, mlp.N_IN))
    lab = np.array([0, 1, 2, 3, 4, 5, 10])
    L, g = mlp.loss_grad(w, feat, lab)
    assert abs(L - np.log(11.0)) < 1e-12
    gb3 = g[mlp.O_B3:mlp.O_MU]
    ref = np.full(mlp.N_OUT, 1.0 / 11.0)
    for y in lab:
        ref[y] -= 1.0 / len(lab)
    assert np.allclose(gb3, ref, atol=1e-12)
    assert np.all(g[: mlp.O_B3] == 0)  # every hidden unit is dead at zero weights
    # the order rule: cumulative probability of the uniform model reaches 0.5 at the 6th order
    o = mlp.select_orders(w, feat, 0.5, 4, 14)
    assert np.all(o == 4 + 5)
    assert np.all(mlp.select_orders(w, feat, 0.5, 10, 14) == 10)
    assert np.all(mlp.select_orders(w, feat, 1.0, 4, 12) == 12)
