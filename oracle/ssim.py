"""CPU restatement of the fused SSIM loss (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
kernels in fastergs_amd/csrc/ssim.hip, never by a product path.

Follows fused-ssim (gs_submodules/fused-ssim/ssim.cu:186-441): 11-tap
Gaussian window, zero padding, separable convolution rows first with the taps
accumulated in order from zero, the SSIM map and its partial derivatives in
the kernel's operation order (float32, no fused multiply-adds), and the
backward as three convolutions of dL/dmap-weighted derivative maps. With
dtype float64 the same formulas give a reference for finite differences.
Pinned by tests/test_ssim_oracle.py (identical images give exactly 1, a
direct 2-D float64 convolution, finite differences of the mean).
"""
from __future__ import annotations

import numpy as np

TAPS = np.array([0.001028380123898387, 0.0075987582094967365, 0.036000773310661316,
                 0.10936068743467331, 0.21300552785396576, 0.26601171493530273,
                 0.21300552785396576, 0.10936068743467331, 0.036000773310661316,
                 0.0075987582094967365, 0.001028380123898387])


def _conv(x, dt):
    """Separable zero-padded 11-tap convolution over the last two axes, taps in order."""
    g = TAPS.astype(dt)
    H, W = x.shape[-2:]
    p = np.zeros(x.shape[:-2] + (H + 10, W + 10), dtype=dt)
    p[..., 5:5 + H, 5:5 + W] = x
    h = np.zeros(x.shape[:-2] + (H + 10, W), dtype=dt)
    for k in range(11):
        h = h + g[k] * p[..., :, k:k + W]
    v = np.zeros(x.shape, dtype=dt)
    for k in range(11):
        v = v + g[k] * h[..., k:k + H, :]
    return v


def forward(img1, img2, C1=0.01 ** 2, C2=0.03 ** 2, dtype=np.float32):
    """(ssim map, dm/dmu1, dm/dsigma1^2, dm/dsigma12), each B x C x H x W."""
    dt = np.dtype(dtype).type
    x1 = np.asarray(img1, dtype=dt)
    x2 = np.asarray(img2, dtype=dt)
    C1, C2, two = dt(C1), dt(C2), dt(2)
    mu1, mu2 = _conv(x1, dt), _conv(x2, dt)
    s1 = _conv(x1 * x1, dt) - mu1 * mu1
    s2 = _conv(x2 * x2, dt) - mu2 * mu2
    s12 = _conv(x1 * x2, dt) - mu1 * mu2
    mu1_sq, mu2_sq, mu1_mu2 = mu1 * mu1, mu2 * mu2, mu1 * mu2
    C = two * mu1_mu2 + C1
    D = two * s12 + C2
    A = mu1_sq + mu2_sq + C1
    B = s1 + s2 + C2
    m = (C * D) / (A * B)
    dmu1 = (mu2 * two * D) / (A * B) - (mu2 * two * C) / (A * B) - \
        (mu1 * two * C * D) / (A * A * B) + (mu1 * two * C * D) / (A * B * B)
    ds1 = (-C * D) / (A * B * B)
    ds12 = (two * C) / (A * B)
    return m, dmu1, ds1, ds12


def backward(img1, img2, dmap, dmu1, ds1, ds12, dtype=np.float32):
    dt = np.dtype(dtype).type
    x1 = np.asarray(img1, dtype=dt)
    x2 = np.asarray(img2, dtype=dt)
    g = np.asarray(dmap, dtype=dt)
    v0 = _conv(np.asarray(dmu1, dtype=dt) * g, dt)
    v1 = _conv(np.asarray(ds1, dtype=dt) * g, dt)
    v2 = _conv(np.asarray(ds12, dtype=dt) * g, dt)
    d = np.zeros_like(x1) + v0
    d = d + x1 * dt(2) * v1
    d = d + x2 * v2
    return d
