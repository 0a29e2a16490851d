"""Fused SSIM on the GPU against the float32 oracle: the map and its three
partial derivatives bit for bit, the backward bit for bit, padding "valid",
train=False, shapes that do not fill the 64 x 16 tiles, and the autograd
entry point (the drop-in ``fused_ssim`` import)."""
import numpy as np
import pytest
import torch

from oracle import ssim as orc

pytestmark = pytest.mark.gpu


def _imgs(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.random(shape).astype(np.float32)
    b = np.clip(a + 0.2 * rng.normal(size=shape), 0, 1).astype(np.float32)
    return a, b


@pytest.mark.parametrize("shape", [(1, 3, 97, 131), (2, 3, 16, 64), (1, 1, 5, 7), (1, 3, 840, 1280)])
def test_maps_and_backward_bit_exact(shape):
    print(f"\n[ssim] {shape}", flush=True)
    from fastergs_amd.ssim import FusedSSIMMap
    a, b = _imgs(shape, shape[2])
    A = torch.tensor(a, device="cuda", requires_grad=True)
    Bt = torch.tensor(b, device="cuda")
    m = FusedSSIMMap.apply(0.01 ** 2, 0.03 ** 2, A, Bt, "same", True)
    rm, rd1, rds1, rds12 = orc.forward(a, b)
    np.testing.assert_array_equal(m.detach().cpu().numpy(), rm)
    g = np.random.default_rng(5).normal(size=shape).astype(np.float32)
    m.backward(torch.tensor(g, device="cuda"))
    rg = orc.backward(a, b, g, rd1, rds1, rds12)
    np.testing.assert_array_equal(A.grad.cpu().numpy(), rg)


def test_valid_padding_and_eval_mode():
    from fused_ssim import fused_ssim
    a, b = _imgs((1, 3, 60, 70), 7)
    A = torch.tensor(a, device="cuda", requires_grad=True)
    Bt = torch.tensor(b, device="cuda")
    loss = fused_ssim(A, Bt, padding="valid")
    rm, rd1, rds1, rds12 = orc.forward(a, b)
    assert abs(loss.item() - float(rm[:, :, 5:-5, 5:-5].astype(np.float64).mean())) < 1e-6
    loss.backward()
    g = np.zeros_like(a)
    g[:, :, 5:-5, 5:-5] = 1.0 / rm[:, :, 5:-5, 5:-5].size
    np.testing.assert_allclose(A.grad.cpu().numpy(), orc.backward(a, b, g.astype(np.float32), rd1,
                                                                  rds1, rds12), rtol=1e-6,
                               atol=1e-12)
    with torch.no_grad():
        v = fused_ssim(A, Bt, train=False)
    assert abs(v.item() - float(rm.astype(np.float64).mean())) < 1e-6
