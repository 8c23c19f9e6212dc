"""FP64 CPU oracle of the library's two networks.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the GPU forward and
backward kernels.  Never part of the product path.

Neither network has a counterpart in the reference (its Newton solves start
from fixed seeds, fem/HexElement.cu:227-272, high_order.h:206, and its iso
search brackets by uniform samples, fem/RayIsosurface.cu:67-188); their
definitions are the ones the library documents:

* FEMRT_MLP_GUESS (``GUESS``): 27 inputs + constant 1 -> 64 tanh -> 64 tanh
  -> 3, loss mean over samples of the sum over outputs of (y - target)^2;
  flat parameters W1 [28][64] (row 27 = first-layer bias), W2 [64][64],
  b2 [64], W3 [64][3], b3 [3].
* FEMRT_MLP_BRACKET (``BRACKET``): 32 inputs -> 64 tanh (bias b1) -> 64 tanh
  -> 2, loss mean over samples of the pinball (quantile) losses of the two
  outputs at quantiles 0.05 and 0.95 against the sample's target (both
  target columns hold the same value); flat parameters W1 [32][64],
  b1 [64], W2 [64][64], b2 [64], W3 [64][2], b3 [2].

``emulate=True`` reproduces the GPU's rounding points (inputs, weights,
hidden activations and back-propagated products rounded to bfloat16 with
round-to-nearest-even; biases and accumulations exact) so the comparison can
be tight; ``emulate=False`` is the exact FP64 function, pinned by finite
differences in tests/test_mlp.py.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

N_H = 64


@dataclass(frozen=True)
class Spec:
    n_in: int
    n_out: int
    bias_input: bool            # first-layer bias as a constant input row
    quantiles: tuple = ()       # pinball loss at these quantiles (else squared error)

    @property
    def w1_rows(self):
        return self.n_in + (1 if self.bias_input else 0)

    @property
    def offsets(self):
        o = {"W1": 0}
        k = self.w1_rows * N_H
        if not self.bias_input:
            o["b1"] = k
            k += N_H
        o["W2"] = k
        k += N_H * N_H
        o["b2"] = k
        k += N_H
        o["W3"] = k
        k += N_H * self.n_out
        o["b3"] = k
        k += self.n_out
        o["end"] = k
        return o

    @property
    def n_param(self):
        return self.offsets["end"]


GUESS = Spec(27, 3, True)
BRACKET = Spec(32, 2, False, (0.05, 0.95))

# the guess network's layout, as module constants
N_IN, N_OUT = GUESS.n_in, GUESS.n_out
OFF_W1 = 0
OFF_W2 = GUESS.offsets["W2"]
OFF_B2 = GUESS.offsets["b2"]
OFF_W3 = GUESS.offsets["W3"]
OFF_B3 = GUESS.offsets["b3"]
N_PARAM = GUESS.n_param


def bf16(x):
    """Round to bfloat16 (nearest, ties to even) through float32; returns
    float64 values."""
    f = np.asarray(x, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(x):
    return np.asarray(x, np.float64).astype(np.float32).astype(np.float64)


def unpack(p, spec=GUESS):
    p = np.asarray(p, np.float64)
    o = spec.offsets
    out = dict(W1=p[o["W1"]:o["W1"] + spec.w1_rows * N_H].reshape(spec.w1_rows, N_H),
               W2=p[o["W2"]:o["b2"]].reshape(N_H, N_H), b2=p[o["b2"]:o["W3"]],
               W3=p[o["W3"]:o["b3"]].reshape(N_H, spec.n_out), b3=p[o["b3"]:o["end"]])
    out["b1"] = p[o["b1"]:o["W2"]] if not spec.bias_input else np.zeros(N_H)
    return out


def pack(W1, W2, b2, W3, b3, b1=None, spec=GUESS):
    parts = [np.ravel(W1)]
    if not spec.bias_input:
        parts.append(np.ravel(b1))
    parts += [np.ravel(W2), np.ravel(b2), np.ravel(W3), np.ravel(b3)]
    return np.concatenate(parts)


def forward(p, feats, emulate=True, spec=GUESS):
    """Outputs n x n_out and the intermediates the backward pass needs."""
    P = unpack(p, spec)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    act = f32 if emulate else (lambda a: a)
    n = feats.shape[0]
    x = np.asarray(feats, np.float64)
    if spec.bias_input:
        x = np.concatenate([x, np.ones((n, 1))], axis=1)
    x = r(x)
    W1, W2, W3 = r(P["W1"]), r(P["W2"]), r(P["W3"])
    b1 = act(P["b1"])
    b2 = act(P["b2"])
    b3 = act(P["b3"])
    h1 = act(np.tanh(x @ W1 + b1))
    h1r = r(h1)
    h2 = act(np.tanh(h1r @ W2 + b2))
    h2r = r(h2)
    y = h2r @ W3 + b3
    return y, dict(x=x, W1=W1, W2=W2, W3=W3, h1=h1, h1r=h1r, h2=h2, h2r=h2r)


def loss_and_output_grad(y, targets, n, spec=GUESS):
    """Loss and dL/dy for outputs y (n x n_out)."""
    t = np.asarray(targets, np.float64)
    if not spec.quantiles:
        diff = y - t
        return float(np.sum(diff ** 2) / n), 2.0 * diff / n
    tau = np.asarray(spec.quantiles, np.float64)[None, :]
    res = t - y
    loss = float(np.sum(np.maximum(tau * res, (tau - 1.0) * res)) / n)
    gy = np.where(res > 0, -tau, 1.0 - tau) / n
    return loss, gy


def loss_grad(p, feats, targets, emulate=True, spec=GUESS):
    """Loss and flat gradient (same layout as the parameters)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    n = feats.shape[0]
    y, c = forward(p, feats, emulate, spec)
    loss, gy = loss_and_output_grad(y, targets, n, spec)
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
    db1 = gz1.sum(0)
    return loss, pack(dW1, dW2, db2, dW3, db3, db1, spec)


def glorot_params(seed, scale=1.0, spec=GUESS):
    """Random parameters for tests (not the library's initialiser)."""
    rng = np.random.default_rng(seed)

    def u(rows, cols, fi, fo):
        a = np.sqrt(6.0 / (fi + fo)) * scale
        return rng.uniform(-a, a, (rows, cols))
    W1 = u(spec.w1_rows, N_H, spec.n_in, N_H)
    W2 = u(N_H, N_H, N_H, N_H)
    W3 = u(N_H, spec.n_out, N_H, spec.n_out)
    b2 = rng.normal(0, 0.1, N_H)
    b3 = rng.normal(0, 0.1, spec.n_out)
    b1 = rng.normal(0, 0.1, N_H)
    return pack(W1, W2, b2, W3, b3, b1, spec)
