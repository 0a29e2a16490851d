"""Pins of the visibility-masked Adam oracle (oracle/optim.py) on the CPU:
the closed form of the first step from zero moments, untouched invisible
rows, and a second step against a float64 restatement."""
import numpy as np

from oracle import optim


def test_first_step_closed_form():
    rng = np.random.default_rng(0)
    N, M = 500, 7
    p = rng.normal(size=(N, M)).astype(np.float32)
    g = (rng.normal(size=(N, M)) * 10.0 ** rng.uniform(-3, 1, (N, M))).astype(np.float32)
    vis = rng.random(N) < 0.6
    lr = 1e-3
    lr_c = optim.compensated_lr(lr, 1)
    # lr_c = lr sqrt(1 - b2) / (1 - b1) at step 1; float32 evaluation of
    # 1 - 0.999 carries a 1.3e-5 relative error (as in the reference)
    assert abs(float(lr_c) - lr * np.sqrt(1 - 0.999) / (1 - 0.9)) <= 2e-5 * float(lr_c)
    p1, m1, v1 = optim.sparse_adam(p, g, np.zeros_like(p), np.zeros_like(p), vis, lr_c)
    # m = (1-b1) g, v = (1-b2) g^2, step = -lr_c m / (sqrt(v) + eps) = -lr sign(g) (|g| >> eps)
    np.testing.assert_allclose(m1[vis], 0.1 * g[vis], rtol=1e-6)
    one_m_b2 = float(np.float32(1) - np.float32(0.999))  # 1 - b2 as float32 forms it
    np.testing.assert_allclose(v1[vis], one_m_b2 * g[vis].astype(np.float64) ** 2, rtol=1e-6)
    np.testing.assert_allclose(p1[vis] - p[vis], -lr * np.sign(g[vis]), rtol=1e-3, atol=1e-7)
    np.testing.assert_array_equal(p1[~vis], p[~vis])
    np.testing.assert_array_equal(m1[~vis], 0)
    np.testing.assert_array_equal(v1[~vis], 0)


def test_second_step_against_float64():
    rng = np.random.default_rng(1)
    N, M = 300, 3
    p = rng.normal(size=(N, M)).astype(np.float32)
    m = rng.normal(size=(N, M)).astype(np.float32) * 0.01
    v = rng.random((N, M)).astype(np.float32) * 0.01
    g = rng.normal(size=(N, M)).astype(np.float32)
    vis = np.ones(N, bool)
    lr = np.float32(2e-3)
    p2, m2, v2 = optim.sparse_adam(p, g, m, v, vis, lr)
    # the decay factors as float32 forms them (1 - 0.999 is 1.3e-5 off in float32)
    b1, b2 = float(np.float32(0.9)), float(np.float32(0.999))
    c1, c2 = float(np.float32(1) - np.float32(0.9)), float(np.float32(1) - np.float32(0.999))
    md = b1 * m.astype(np.float64) + c1 * g
    vd = b2 * v.astype(np.float64) + c2 * g.astype(np.float64) ** 2
    pd = p - float(lr) * md / (np.sqrt(vd) + 1e-15)
    np.testing.assert_allclose(m2, md, rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(v2, vd, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(p2, pd, rtol=1e-6, atol=1e-7)
