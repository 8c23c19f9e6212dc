"""FP64 CPU oracle of the learned Newton-guess network.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the GPU forward and
backward kernels.  Never part of the product path.

The network has no counterpart in the reference (its Newton solves start
from fixed seeds: fem/HexElement.cu:227-272, high_order.h:206); its definition
is the one the library documents for FEMRT_MLP_GUESS: 27 inputs + constant 1
-> 64 tanh -> 64 tanh -> 3, loss mean over samples of sum over outputs of
(y - target)^2, flat parameters W1 [28][64] (row 27 = first-layer bias),
W2 [64][64], b2 [64], W3 [64][3], b3 [3].

``emulate=True`` reproduces the GPU's rounding points (inputs, weights,
hidden activations and back-propagated products rounded to bfloat16 with
round-to-nearest-even; biases and accumulations exact) so the comparison can
be tight; ``emulate=False`` is the exact FP64 function, pinned by finite
differences in tests/test_mlp.py.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 27, 64, 3
OFF_W1 = 0
OFF_W2 = 28 * 64
OFF_B2 = OFF_W2 + 64 * 64
OFF_W3 = OFF_B2 + 64
OFF_B3 = OFF_W3 + 64 * 3
N_PARAM = OFF_B3 + 3


def bf16(x):
    """Round to bfloat16 (nearest, ties to even) through float32; returns
    float64 values."""
    f = np.asarray(x, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(x):
    return np.asarray(x, np.float64).astype(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return dict(W1=p[OFF_W1:OFF_W2].reshape(28, 64), W2=p[OFF_W2:OFF_B2].reshape(64, 64),
                b2=p[OFF_B2:OFF_W3], W3=p[OFF_W3:OFF_B3].reshape(64, 3), b3=p[OFF_B3:N_PARAM])


def pack(W1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(W1), np.ravel(W2), np.ravel(b2), np.ravel(W3), np.ravel(b3)])


def forward(p, feats, emulate=True):
    """Outputs n x 3 and the intermediates the backward pass needs."""
    P = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    act = f32 if emulate else (lambda a: a)
    n = feats.shape[0]
    x = r(np.concatenate([np.asarray(feats, np.float64), np.ones((n, 1))], axis=1))
    W1, W2, W3 = r(P["W1"]), r(P["W2"]), r(P["W3"])
    b2 = act(P["b2"])
    b3 = act(P["b3"])
    h1 = act(np.tanh(x @ W1))
    h1r = r(h1)
    h2 = act(np.tanh(h1r @ W2 + b2))
    h2r = r(h2)
    y = h2r @ W3 + b3
    return y, dict(x=x, W1=W1, W2=W2, W3=W3, h1=h1, h1r=h1r, h2=h2, h2r=h2r)


def loss_grad(p, feats, targets, emulate=True):
    """Loss and flat gradient (same layout as the parameters)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    n = feats.shape[0]
    y, c = forward(p, feats, emulate)
    diff = y - np.asarray(targets, np.float64)
    loss = float(np.sum(diff ** 2) / n)
    gy = 2.0 * diff / n
    gyr = r(gy)
    dW3 = c["h2r"].T @ gyr
    db3 = gy.sum(0)
    gz2 = (gyr @ c["W3"].T) * (1.0 - c["h2"] ** 2)
    gz2r = r(gz2)
    db2 = gz2.sum(0)
    dW2 = c["h1r"].T @ gz2r
    gz1 = (gz2r @ c["W2"].T) * (1.0 - c["h1"] ** 2)
    gz1r = r(gz1)
    dW1 = c["x"].T @ gz1r
    return loss, pack(dW1, dW2, db2, dW3, db3)


def glorot_params(seed, scale=1.0):
    """Random parameters for tests (not the library's initialiser)."""
    rng = np.random.default_rng(seed)

    def u(rows, cols, fi, fo):
        a = np.sqrt(6.0 / (fi + fo)) * scale
        return rng.uniform(-a, a, (rows, cols))
    W1 = u(28, 64, 27, 64)
    W2 = u(64, 64, 64, 64)
    W3 = u(64, 3, 64, 3)
    b2 = rng.normal(0, 0.1, 64)
    b3 = rng.normal(0, 0.1, 3)
    return pack(W1, W2, b2, W3, b3)
