"""Pins of the SSIM oracle (oracle/ssim.py) on the CPU: identical images give
exactly 1; the separable convolution equals a direct 2-D float64 one; the
analytic backward matches float64 finite differences of the mean SSIM."""
import numpy as np

from oracle import ssim


def _imgs(shape=(2, 3, 23, 37), seed=0):
    rng = np.random.default_rng(seed)
    a = rng.random(shape).astype(np.float32)
    b = np.clip(a + 0.1 * rng.normal(size=shape), 0, 1).astype(np.float32)
    return a, b


def test_identical_images_give_one():
    a, _ = _imgs()
    m, *_ = ssim.forward(a, a)
    np.testing.assert_array_equal(m, np.ones_like(m))


def test_separable_equals_direct_2d():
    a, b = _imgs((1, 1, 19, 21), 1)
    w = np.outer(ssim.TAPS, ssim.TAPS)
    H, W = a.shape[-2:]
    pa = np.pad(a[0, 0].astype(np.float64), 5)
    direct = np.array([[np.sum(w * pa[y:y + 11, x:x + 11]) for x in range(W)] for y in range(H)])
    np.testing.assert_allclose(ssim._conv(a.astype(np.float64), np.float64)[0, 0], direct,
                               rtol=1e-12, atol=1e-14)
    m64, *_ = ssim.forward(a, b, dtype=np.float64)
    m32, *_ = ssim.forward(a, b)
    np.testing.assert_allclose(m32, m64, atol=2e-5)


def test_backward_matches_finite_differences():
    a, b = _imgs((1, 2, 12, 14), 2)
    a64 = a.astype(np.float64)
    m, d1, ds1, ds12 = ssim.forward(a64, b, dtype=np.float64)
    dmap = np.full_like(m, 1.0 / m.size)  # d mean / d map
    grad = ssim.backward(a64, b, dmap, d1, ds1, ds12, dtype=np.float64)
    rng = np.random.default_rng(3)
    for _ in range(40):
        idx = tuple(rng.integers(0, s) for s in a.shape)
        h = 1e-6
        ap, am = a64.copy(), a64.copy()
        ap[idx] += h
        am[idx] -= h
        fd = (ssim.forward(ap, b, dtype=np.float64)[0].mean() -
              ssim.forward(am, b, dtype=np.float64)[0].mean()) / (2 * h)
        assert abs(fd - grad[idx]) <= 1e-6 + 1e-5 * abs(fd), (idx, fd, grad[idx])
