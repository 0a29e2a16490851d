"""CPU restatement of the tile cost model (numpy, float64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
cost-model kernels (fastergs_amd/csrc/sched.hip), never by a product path.

The model has no counterpart in the FasterGS reference (its tile schedule is
fixed); it is defined by its mathematics:
    h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3
    loss = mean_r [ (y0 - t0)^2 + s * BCE(sigmoid(y1), t1) ]
with the parameter layout W1 [64][8], b1 [64], W2 [64][64], b2 [64],
W3 [2][64], b3 [2] (4866 floats). ``bf16=True`` rounds weights, inputs and
hidden activations to bfloat16 (round to nearest even) as the inference
kernel does. Pinned by central finite differences and by an independent
torch autograd restatement (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 8, 64, 2
SHAPES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (OUT, HID)), ("b3", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(params: np.ndarray) -> dict:
    p = np.asarray(params, dtype=np.float64).ravel()
    assert p.size == NPARAMS
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(params, x, bf16: bool = False) -> np.ndarray:
    w = unpack(params)
    x = np.asarray(x, dtype=np.float64).reshape(-1, IN)
    if bf16:
        w = {k: bf16_round(v.astype(np.float32)) for k, v in w.items()}
        x = bf16_round(x.astype(np.float32))
    h1 = np.maximum(x @ w["W1"].T + w["b1"], 0.0)
    if bf16:
        h1 = bf16_round(h1.astype(np.float32))
    h2 = np.maximum(h1 @ w["W2"].T + w["b2"], 0.0)
    if bf16:
        h2 = bf16_round(h2.astype(np.float32))
    return h2 @ w["W3"].T + w["b3"]


def loss_grad(params, x, target, split_weight: float = 1.0):
    """Mean loss and its gradient w.r.t. the flat parameter vector."""
    w = unpack(params)
    x = np.asarray(x, dtype=np.float64).reshape(-1, IN)
    t = np.asarray(target, dtype=np.float64).reshape(-1, OUT)
    n = x.shape[0]
    z1 = x @ w["W1"].T + w["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w["W3"].T + w["b3"]
    d0 = y[:, 0] - t[:, 0]
    y1 = y[:, 1]
    bce = np.maximum(y1, 0.0) - y1 * t[:, 1] + np.log1p(np.exp(-np.abs(y1)))
    loss = float(np.mean(d0 * d0 + split_weight * bce))
    gy = np.stack([2.0 * d0, split_weight * (1.0 / (1.0 + np.exp(-y1)) - t[:, 1])], axis=1) / n
    g = {"W3": gy.T @ h2, "b3": gy.sum(0)}
    gz2 = (gy @ w["W3"]) * (z2 > 0)
    g["W2"], g["b2"] = gz2.T @ h1, gz2.sum(0)
    gz1 = (gz2 @ w["W2"]) * (z1 > 0)
    g["W1"], g["b1"] = gz1.T @ x, gz1.sum(0)
    return loss, np.concatenate([g[name].ravel() for name, _ in SHAPES])


def adam(params, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (torch.optim.Adam, no weight decay)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
