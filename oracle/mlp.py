"""CPU restatement of the surrogate visibility network (numpy).

TEST INFRASTRUCTURE ONLY: the checker that the HIP forward and training
kernels are compared against.  The network has no counterpart in the
reference; it stands in for the cone trace of
shader/cacheLightingRSM.comp:167-231.

Network: 32 -> 64 -> 64 -> 64 -> 1, ReLU hidden layers, sigmoid output,
binary cross-entropy loss averaged over the batch.  Parameters are one FP32
vector in the layout W1[32][64] (row = input), b1[64], W2[64][64], b2,
W3[64][64], b3, w4[64], b4.

``bf16=True`` reproduces where the matrix-core kernels round to bfloat16
(round to nearest even): the network inputs, W1..W3, every hidden activation
after ReLU, and the backpropagated pre-activation gradients that feed the
gradient GEMMs.  Everything else (biases, w4, the output unit, bias
gradients, accumulation) is exact.  ``bf16=False`` is the plain float64
network whose gradient the finite-difference tests pin.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 32, 64
W1, B1 = 0, 2048
W2, B2 = 2112, 6208
W3, B3 = 6272, 10368
W4, B4 = 10432, 10496
N_PARAMS = 10497


def bf16_round(a: np.ndarray) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    rounded = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = rounded.astype(np.uint32).view(np.float32)
    # NaN/Inf pass through unchanged
    special = ~np.isfinite(f)
    out = np.where(special, f, out)
    return out.astype(np.float64)


def unpack(params: np.ndarray) -> dict[str, np.ndarray]:
    p = np.asarray(params, dtype=np.float64)
    assert p.shape == (N_PARAMS,)
    return {
        "W1": p[W1:B1].reshape(N_IN, N_HID), "b1": p[B1:W2],
        "W2": p[W2:B2].reshape(N_HID, N_HID), "b2": p[B2:W3],
        "W3": p[W3:B3].reshape(N_HID, N_HID), "b3": p[B3:W4],
        "w4": p[W4:B4], "b4": p[B4],
    }


def pack(d: dict[str, np.ndarray]) -> np.ndarray:
    out = np.zeros(N_PARAMS, dtype=np.float64)
    out[W1:B1] = d["W1"].ravel()
    out[B1:W2] = d["b1"]
    out[W2:B2] = d["W2"].ravel()
    out[B2:W3] = d["b2"]
    out[W3:B3] = d["W3"].ravel()
    out[B3:W4] = d["b3"]
    out[W4:B4] = d["w4"]
    out[B4] = d["b4"]
    return out


def _r(a, on):
    return bf16_round(a) if on else np.asarray(a, dtype=np.float64)


def forward(params: np.ndarray, x: np.ndarray, bf16: bool = True):
    """Returns (p, cache) for features x[n][32]."""
    d = unpack(params)
    xb = _r(x, bf16)
    h1 = _r(np.maximum(xb @ _r(d["W1"], bf16) + d["b1"], 0.0), bf16)
    h2 = _r(np.maximum(h1 @ _r(d["W2"], bf16) + d["b2"], 0.0), bf16)
    h3 = _r(np.maximum(h2 @ _r(d["W3"], bf16) + d["b3"], 0.0), bf16)
    z = h3 @ d["w4"] + d["b4"]
    p = 1.0 / (1.0 + np.exp(-z))
    return p, {"x": xb, "h1": h1, "h2": h2, "h3": h3, "z": z, "d": d}


def loss_and_grad(params: np.ndarray, x: np.ndarray, y: np.ndarray, bf16: bool = True):
    """Summed BCE loss and the gradient of the MEAN loss over the batch."""
    n = x.shape[0]
    p, c = forward(params, x, bf16)
    d = c["d"]
    y = np.asarray(y, dtype=np.float64)
    pc = np.clip(p, 1e-7, 1.0 - 1e-7)
    loss = float(-np.sum(y * np.log(pc) + (1.0 - y) * np.log(1.0 - pc)))
    dz4 = (p - y) / n
    g = {}
    g["w4"] = c["h3"].T @ dz4
    g["b4"] = np.sum(dz4)
    dz3 = np.where(c["h3"] > 0.0, np.outer(dz4, d["w4"]), 0.0)
    g["b3"] = dz3.sum(0)
    dz3r = _r(dz3, bf16)
    g["W3"] = c["h2"].T @ dz3r
    dz2 = np.where(c["h2"] > 0.0, dz3r @ _r(d["W3"], bf16).T, 0.0)
    g["b2"] = dz2.sum(0)
    dz2r = _r(dz2, bf16)
    g["W2"] = c["h1"].T @ dz2r
    dz1 = np.where(c["h1"] > 0.0, dz2r @ _r(d["W2"], bf16).T, 0.0)
    g["b1"] = dz1.sum(0)
    g["W1"] = c["x"].T @ _r(dz1, bf16)
    return loss, pack(g)


def adam_step(params, grad, m, v, t: int, lr: float, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t >= 1) on float64 copies; returns (params, m, v)."""
    m = b1 * m + (1.0 - b1) * grad
    v = b2 * v + (1.0 - b2) * grad * grad
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
