"""CPU restatement of the visibility-masked Adam update (numpy, float32).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
kernel `k_sparse_adam`, never by a product path.

Follows adamUpdateCUDA (FGS/csrc/cuda_rasterizer/adam.cu:9-38) operation by
operation in float32 (no fused multiply-adds), and SparseGaussianAdam.step
(FGS/diff_gaussian_rasterization_faster/__init__.py:267-298) for the step
size compensation. Pinned by the closed form of the first step
(tests/test_optim_oracle.py).
"""
from __future__ import annotations

import numpy as np


def compensated_lr(lr: float, step: int) -> np.float32:
    """lr * sqrt(1 - 0.999^step) / (1 - 0.9^step), evaluated like the
    reference (float32 tensor arithmetic on a float32 step counter)."""
    s = np.float32(step)
    return np.float32(np.float32(lr) * np.sqrt(np.float32(1) - np.float32(0.999) ** s)
                      / (np.float32(1) - np.float32(0.9) ** s))


def sparse_adam(param, grad, m, v, visible, lr, b1=0.9, b2=0.999, eps=1e-15):
    """One update of N x M float32 arrays in place-free form; returns (p, m, v)."""
    p = np.array(param, dtype=np.float32, copy=True)
    m = np.array(m, dtype=np.float32, copy=True)
    v = np.array(v, dtype=np.float32, copy=True)
    g = np.asarray(grad, dtype=np.float32)
    N = np.asarray(visible).shape[0]
    p2, m2, v2, g2 = p.reshape(N, -1), m.reshape(N, -1), v.reshape(N, -1), g.reshape(N, -1)
    vis = np.asarray(visible).astype(bool)
    f32 = np.float32
    b1, b2, lr, eps = f32(b1), f32(b2), f32(lr), f32(eps)
    mm = b1 * m2[vis] + (f32(1) - b1) * g2[vis]
    vv = b2 * v2[vis] + ((f32(1) - b2) * g2[vis]) * g2[vis]
    step = (-lr * mm) / (np.sqrt(vv) + eps)
    p2[vis] = p2[vis] + step
    m2[vis] = mm
    v2[vis] = vv
    return p, m, v
