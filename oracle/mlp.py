"""CPU restatement of the order-selection MLP of libnbfmm (forward pass,
mean cross entropy and its gradient), in float64 numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.

The model has no counterpart in the reference (one expansion order L for all
cells, tree.cpp:94), so nothing in the reference pins it; this restatement is
pinned by central finite differences of its own loss (tests/test_oracle_mlp.py)
and then used to check the GPU forward, backward and weight-gradient kernels.

Weight vector (5995 float32): W1[16][64], b1[64], W2[64][64], b2[64],
W3[64][11], b3[11] (the 5963 trainable values), mean[16], deviation[16].
Forward: x = (f - mean) / deviation; a1 = relu(x W1 + b1);
a2 = relu(a1 W2 + b2); z = a2 W3 + b3; p = softmax(z).  The GPU uses the
weight matrices rounded to bfloat16 (round to nearest even); `bf16=True`
does the same here.
"""
import numpy as np

N_IN, N_H, N_OUT = 16, 64, 11
O_W1 = 0
O_B1 = O_W1 + N_IN * N_H
O_W2 = O_B1 + N_H
O_B2 = O_W2 + N_H * N_H
O_W3 = O_B2 + N_H
O_B3 = O_W3 + N_H * N_OUT
O_MU = O_B3 + N_OUT
O_SD = O_MU + N_IN
N_PARAMS = O_MU
N_WEIGHTS = O_SD + N_IN


def bf16_round(a):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def unpack(w, bf16=True):
    w = np.asarray(w, dtype=np.float32)
    r = (lambda a: bf16_round(a)) if bf16 else (lambda a: a)
    return {
        "W1": r(w[O_W1:O_B1]).astype(np.float64).reshape(N_IN, N_H),
        "b1": w[O_B1:O_W2].astype(np.float64),
        "W2": r(w[O_W2:O_B2]).astype(np.float64).reshape(N_H, N_H),
        "b2": w[O_B2:O_W3].astype(np.float64),
        "W3": r(w[O_W3:O_B3]).astype(np.float64).reshape(N_H, N_OUT),
        "b3": w[O_B3:O_MU].astype(np.float64),
        "mu": w[O_MU:O_SD].astype(np.float64),
        "sd": w[O_SD:N_WEIGHTS].astype(np.float64),
    }


def forward(w, feat, bf16=True):
    P = unpack(w, bf16) if not isinstance(w, dict) else w
    x = (np.asarray(feat, np.float64) - P["mu"]) / P["sd"]
    a1 = np.maximum(x @ P["W1"] + P["b1"], 0.0)
    a2 = np.maximum(a1 @ P["W2"] + P["b2"], 0.0)
    z = a2 @ P["W3"] + P["b3"]
    return x, a1, a2, z


def softmax(z):
    e = np.exp(z - z.max(axis=1, keepdims=True))
    return e / e.sum(axis=1, keepdims=True)


def probs(w, feat, bf16=True):
    return softmax(forward(w, feat, bf16)[3])


def loss(w, feat, label, bf16=True):
    p = probs(w, feat, bf16)
    return float(-np.log(p[np.arange(len(label)), np.asarray(label, np.int64)]).mean())


def loss_grad(w, feat, label, bf16=True):
    """(mean cross entropy, gradient over the N_PARAMS trainable values)."""
    P = unpack(w, bf16)
    x, a1, a2, z = forward(P, feat)
    p = softmax(z)
    y = np.asarray(label, np.int64)
    B = len(y)
    L = float(-np.log(p[np.arange(B), y]).mean())
    d3 = p.copy()
    d3[np.arange(B), y] -= 1.0
    d3 /= B
    d2 = (d3 @ P["W3"].T) * (a2 > 0)
    d1 = (d2 @ P["W2"].T) * (a1 > 0)
    g = np.zeros(N_PARAMS)
    g[O_W1:O_B1] = (x.T @ d1).ravel()
    g[O_B1:O_W2] = d1.sum(0)
    g[O_W2:O_B2] = (a1.T @ d2).ravel()
    g[O_B2:O_W3] = d2.sum(0)
    g[O_W3:O_B3] = (a2.T @ d3).ravel()
    g[O_B3:O_MU] = d3.sum(0)
    return L, g


def select_orders(w, feat, conf, omin, omax, lo=4):
    """Lowest order whose cumulative probability reaches conf, clamped to [omin, omax]."""
    c = np.cumsum(probs(w, feat), axis=1)
    k = np.argmax(c >= conf, axis=1)
    k[~(c >= conf).any(axis=1)] = N_OUT - 1
    return np.clip(lo + k, omin, omax)
