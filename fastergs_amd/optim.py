"""SparseGaussianAdam: the visibility-masked Adam of the FasterGS training
loops (diff_gaussian_rasterization_faster/__init__.py:267-298), e.g.
mini-splatting2/msv2/train_faster.py:248:

    optimizer = SparseGaussianAdam(param_groups, lr=0.0, eps=1e-15)
    ...
    loss.backward()
    optimizer.step(visibility=radii > 0, N=num_gaussians)

Same constructor, ``step(visibility, N)`` signature, lazily created state
(``step`` as a float32 tensor, ``exp_avg``, ``exp_avg_sq``), one tensor per
group and the same bias compensation of the step size,
lr * sqrt(1 - 0.999^step) / (1 - 0.9^step). Difference: all groups of a step
are updated by one kernel launch (``fgs_sparse_adam``) instead of one per
group.
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import AdamTensor, check, lib


class SparseGaussianAdam(torch.optim.Adam):
    def __init__(self, params, lr, eps):
        super().__init__(params=params, lr=lr, eps=eps)

    @torch.no_grad()
    def step(self, visibility, N):
        items = []
        dev = None
        for group in self.param_groups:
            lr = group["lr"]
            eps = group["eps"]
            assert len(group["params"]) == 1, "more than one tensor in group"
            param = group["params"][0]
            if param.grad is None or param.grad.numel() == 0:
                continue
            state = self.state[param]
            if len(state) == 0:
                state["step"] = torch.tensor(0.0, dtype=torch.float32)
                state["exp_avg"] = torch.zeros_like(param, memory_format=torch.preserve_format)
                state["exp_avg_sq"] = torch.zeros_like(param, memory_format=torch.preserve_format)
            state["step"] += 1
            step = state["step"]
            lr_c = float(lr * (1 - 0.999 ** step) ** 0.5 / (1 - 0.9 ** step))
            if param.dtype != torch.float32 or not param.is_contiguous():
                raise ValueError("SparseGaussianAdam needs contiguous float32 parameters")
            grad = param.grad.contiguous().to(torch.float32)
            M = param.numel() // N
            items.append((param, grad, state["exp_avg"], state["exp_avg_sq"], M, lr_c, eps))
            dev = param.device
        if not items:
            return
        vis = visibility.to(device=dev, dtype=torch.uint8).contiguous()
        arr = (AdamTensor * len(items))()
        for k, (p, g, m, v, M, lr_c, eps) in enumerate(items):
            arr[k] = AdamTensor(p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), M, lr_c,
                                float(eps))
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        for start in range(0, len(items), 8):  # the library takes 8 tensors per launch
            n = min(8, len(items) - start)
            sub = ctypes.cast(ctypes.byref(arr, start * ctypes.sizeof(AdamTensor)),
                              ctypes.POINTER(AdamTensor))
            check(lib().fgs_sparse_adam(sub, n, ctypes.c_void_p(vis.data_ptr()), int(N), 0.9,
                                        0.999, stream), "fgs_sparse_adam")
