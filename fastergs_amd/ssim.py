"""Fused SSIM loss on gfx950, drop-in for the ``fused_ssim`` package the
FasterGS training loops import (FastGS/train_faster.py:18,
gs_submodules/fused-ssim/fused_ssim/__init__.py):

    from fused_ssim import fused_ssim
    loss = (1 - lambda_) * l1 + lambda_ * (1.0 - fused_ssim(image[None], gt[None]))

Same signature (``img1, img2, padding="same" | "valid", train=True``), same
constants C1 = 0.01^2, C2 = 0.03^2, mean of the map, gradient w.r.t. img1
only. The map, its partial derivatives and the backward are HIP kernels of
libfastergs.so (``fgs_ssim_forward`` / ``fgs_ssim_backward``).
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import check, lib

allowed_padding = ["same", "valid"]


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else None


def _s(t):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


class FusedSSIMMap(torch.autograd.Function):
    @staticmethod
    def forward(ctx, C1, C2, img1, img2, padding="same", train=True):
        img1c = img1.detach().to(torch.float32).contiguous()
        img2c = img2.to(torch.float32).contiguous()
        if img1c.dim() != 4 or img1c.shape != img2c.shape:
            raise ValueError("fused_ssim expects two B x C x H x W images of the same shape")
        B, CH, H, W = img1c.shape
        ssim_map = torch.empty_like(img1c)
        derivs = [torch.empty_like(img1c) for _ in range(3)] if train else [None] * 3
        check(lib().fgs_ssim_forward(float(C1), float(C2), _p(img1c), _p(img2c), B, CH, H, W,
                                     _p(ssim_map), *[_p(d) for d in derivs], _s(img1c)),
              "fgs_ssim_forward")
        if padding == "valid":
            ssim_map = ssim_map[:, :, 5:-5, 5:-5]
        ctx.save_for_backward(img1c, img2c, *(d for d in derivs if d is not None))
        ctx.padding = padding
        ctx.train = train
        return ssim_map

    @staticmethod
    def backward(ctx, opt_grad):
        if not ctx.train:
            raise RuntimeError("fused_ssim was evaluated with train=False: no gradient")
        img1, img2, dmu1, ds1, ds12 = ctx.saved_tensors
        dL_dmap = opt_grad
        if ctx.padding == "valid":
            dL_dmap = torch.zeros_like(img1)
            dL_dmap[:, :, 5:-5, 5:-5] = opt_grad
        dL_dmap = dL_dmap.to(torch.float32).contiguous()
        grad = torch.empty_like(img1)
        B, CH, H, W = img1.shape
        check(lib().fgs_ssim_backward(_p(img1), _p(img2), _p(dL_dmap), _p(dmu1), _p(ds1),
                                      _p(ds12), B, CH, H, W, _p(grad), _s(img1)),
              "fgs_ssim_backward")
        return None, None, grad, None, None, None


def fused_ssim(img1, img2, padding="same", train=True):
    C1 = 0.01 ** 2
    C2 = 0.03 ** 2
    assert padding in allowed_padding
    ssim_map = FusedSSIMMap.apply(C1, C2, img1, img2, padding, train)
    return ssim_map.mean()
