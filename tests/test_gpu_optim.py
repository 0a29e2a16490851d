"""SparseGaussianAdam on the GPU against the float32 oracle (bit-exact):
the six parameter groups of a 3DGS model (means 3, dc 3, rest SH 45,
opacity 1, scales 3, rotations 4 per Gaussian), random visibility, three
steps; more than 8 groups (several launches); state layout as the
reference's (step tensor, exp_avg, exp_avg_sq)."""
import numpy as np
import pytest
import torch

from fastergs_amd.optim import SparseGaussianAdam
from oracle import optim

pytestmark = pytest.mark.gpu

SHAPES = {"xyz": (3,), "f_dc": (1, 3), "f_rest": (15, 3), "opacity": (1,), "scaling": (3,),
          "rotation": (4,)}


def _lr_c(lr, step):  # the optimizer's own compensation (float32 torch arithmetic)
    s = torch.tensor(float(step), dtype=torch.float32)
    return float(lr * (1 - 0.999 ** s) ** 0.5 / (1 - 0.9 ** s))


@pytest.mark.parametrize("N,groups", [(10007, SHAPES), (513, {f"g{k}": (k % 5 + 1,) for k in range(11)})])
def test_sparse_adam_matches_oracle(N, groups):
    print(f"\n[adam] N={N} groups={len(groups)}", flush=True)
    rng = np.random.default_rng(N)
    params = {k: torch.tensor(rng.normal(size=(N,) + s).astype(np.float32), device="cuda",
                              requires_grad=True) for k, s in groups.items()}
    lrs = {k: float(10.0 ** rng.uniform(-4, -2)) for k in groups}
    opt = SparseGaussianAdam([{"params": [p], "lr": lrs[k], "name": k} for k, p in params.items()],
                             lr=0.0, eps=1e-15)
    ref = {k: (p.detach().cpu().numpy().copy(), np.zeros((N,) + groups[k], np.float32),
               np.zeros((N,) + groups[k], np.float32)) for k, p in params.items()}
    for step in range(1, 4):
        vis = rng.random(N) < 0.55
        for k, p in params.items():
            p.grad = torch.tensor(rng.normal(size=p.shape).astype(np.float32), device="cuda")
        grads = {k: p.grad.cpu().numpy() for k, p in params.items()}
        opt.step(torch.tensor(vis, device="cuda"), N)
        for k, p in params.items():
            P0, m0, v0 = ref[k]
            ref[k] = optim.sparse_adam(P0, grads[k], m0, v0, vis, _lr_c(lrs[k], step))
            st = opt.state[p]
            assert float(st["step"]) == step
            np.testing.assert_array_equal(p.detach().cpu().numpy(), ref[k][0], err_msg=k)
            np.testing.assert_array_equal(st["exp_avg"].cpu().numpy(), ref[k][1], err_msg=k)
            np.testing.assert_array_equal(st["exp_avg_sq"].cpu().numpy(), ref[k][2], err_msg=k)
