"""Pins of the cost-model oracle (oracle/mlp.py) on the CPU.

* loss gradient against float64 central finite differences;
* forward and gradient against an independent torch float64 autograd
  restatement written from the model's definition;
* bfloat16 rounding on hand-picked values (ties to even, carries into the
  exponent, negative values);
* the Adam step against torch.optim.Adam.
"""
import numpy as np
import torch

from oracle import mlp


def _params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, scale, mlp.NPARAMS)


def _data(n=37, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, 1.0, (n, mlp.IN))
    t = np.stack([rng.normal(6.0, 2.0, n), (rng.random(n) < 0.3).astype(np.float64)], 1)
    return x, t


def test_gradient_matches_finite_differences():
    p = _params()
    x, t = _data()
    _, g = mlp.loss_grad(p, x, t, split_weight=0.7)
    rng = np.random.default_rng(5)
    idx = rng.choice(mlp.NPARAMS, 200, replace=False)
    idx = np.concatenate([idx, [0, 511, 512, 575, 576, 4671, 4672, 4735, 4736, 4863, 4864, 4865]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss_grad(pp, x, t, 0.7)[0] - mlp.loss_grad(pm, x, t, 0.7)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (1 + abs(fd)), (i, fd, g[i])


def test_matches_torch_autograd():
    p = _params(3)
    x, t = _data(50, 4)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    o = 0
    W = {}
    for name, shape in mlp.SHAPES:
        n = int(np.prod(shape))
        W[name] = pt[o:o + n].reshape(shape)
        o += n
    xt = torch.tensor(x)
    tt = torch.tensor(t)
    h1 = torch.relu(xt @ W["W1"].T + W["b1"])
    h2 = torch.relu(h1 @ W["W2"].T + W["b2"])
    y = h2 @ W["W3"].T + W["b3"]
    bce = torch.nn.functional.binary_cross_entropy_with_logits(y[:, 1], tt[:, 1], reduction="none")
    loss = ((y[:, 0] - tt[:, 0]) ** 2 + 1.3 * bce).mean()
    loss.backward()
    l_ref, g_ref = mlp.loss_grad(p, x, t, 1.3)
    np.testing.assert_allclose(mlp.forward(p, x), y.detach().numpy(), rtol=1e-12, atol=1e-12)
    assert abs(l_ref - loss.item()) < 1e-12
    np.testing.assert_allclose(g_ref, pt.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20,
                  -1.0 - 2 ** -8 - 2 ** -20, 1.9999999, 0.0], dtype=np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -1.0 - 2 ** -7, 2.0, 0.0])
    np.testing.assert_array_equal(mlp.bf16_round(v), want)
    big = mlp.bf16_round(np.array([3.0e38], dtype=np.float32)).astype(np.float32)
    assert int(big.view(np.uint32)[0]) & 0xFFFF == 0 and abs(big[0] / 3.0e38 - 1) < 2 ** -8


def test_bf16_forward_differs_only_by_rounding():
    p = _params(7)
    x, _ = _data(64, 8)
    y32 = mlp.forward(p, x)
    y16 = mlp.forward(p, x, bf16=True)
    rel = np.abs(y16 - y32).max() / (np.abs(y32).max() + 1e-30)
    assert 1e-5 < rel < 5e-2


def test_adam_matches_torch():
    rng = np.random.default_rng(9)
    p0 = rng.normal(size=100)
    pt = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(100), np.zeros(100)
    for step in range(1, 6):
        g = rng.normal(size=100)
        pt.grad = torch.tensor(g)
        opt.step()
        p, m, v = mlp.adam(p, g, m, v, step, lr=1e-2)
    np.testing.assert_allclose(p, pt.detach().numpy(), rtol=1e-12, atol=1e-12)
