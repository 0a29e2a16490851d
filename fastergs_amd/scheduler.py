"""Tile schedules of the FAST compositing pass and the learned cost model.

The compositing pass runs one workgroup per work item: a 16x16 tile or one
of its two 16x8 row bands. The schedule decides which tiles are split and in
which order items are dispatched; the rendered outputs and the gradients
never depend on it (tests/test_gpu_sched.py checks them bit for bit), only
the kernel time does.

* ``TileSchedule.raster()``: one item per tile in tile order (the
  reference's launch, render.cu:418).
* ``TileSchedule.heuristic(split_len)``: tile order, tiles with at least
  ``split_len`` list entries split -- a fixed threshold, the fallback.
* ``TileSchedule.learned(model, order=False)``: a cost model predicts each
  tile's log2 duration and a split flag from 8 features of the tile's list;
  the split flags are applied in tile order, or with ``order=True`` the
  items are dispatched longest-predicted first. With ``ahead=True`` the
  longest-predicted-first list of a frame is built from the previous
  frame's tile features on a side stream beside its preprocess, and those
  features are computed beside the previous frame's compositing, so no
  scheduling kernel runs on the critical path (on a camera trajectory
  consecutive views are alike).

``CostModel`` is the MLP (8 -> 64 -> 64 -> 2, fp32 weights, bf16
inference). Its forward, loss gradient and Adam update are HIP kernels of
libfastergs.so. ``OnlineTrainer`` trains it on the durations the
compositing pass logs for every item of the frames it is shown.
"""
from __future__ import annotations

import ctypes
import math

import torch

from . import _lib
from ._lib import TileSched, check, lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else None


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


class CostModel:
    """Per-tile cost model on the device. Outputs per tile: predicted log2
    of the compositing duration in 100 MHz ticks, and a split logit."""

    def __init__(self, device="cuda", seed: int = 0, lr: float = 1e-3):
        self.device = torch.device(device)
        g = torch.Generator().manual_seed(seed)
        parts = []
        for fan_in, fan_out in ((_lib.MLP_IN, _lib.MLP_HIDDEN), (_lib.MLP_HIDDEN, _lib.MLP_HIDDEN),
                                (_lib.MLP_HIDDEN, _lib.MLP_OUT)):
            bound = math.sqrt(6.0 / fan_in)  # He-uniform for ReLU layers
            parts.append((torch.rand(fan_out * fan_in, generator=g) * 2 - 1) * bound)
            parts.append(torch.zeros(fan_out))
        parts[-2] *= 0.1
        parts[-1] = torch.tensor([8.0, 0.0])  # start near a typical log2 duration
        self.params = torch.cat(parts).to(self.device, torch.float32).contiguous()
        assert self.params.numel() == _lib.MLP_NPARAMS
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.loss = torch.zeros(1, device=self.device)
        self.lr = lr
        self.steps = 0
        self._ws = None

    def forward(self, x: torch.Tensor, bf16: bool = True) -> torch.Tensor:
        x = x.to(self.device, torch.float32).contiguous().reshape(-1, _lib.MLP_IN)
        y = torch.empty(x.shape[0], _lib.MLP_OUT, device=self.device)
        check(lib().fgs_mlp_forward(_p(self.params), _p(x), x.shape[0], _p(y), 1 if bf16 else 0,
                                    _stream(self.device)), "fgs_mlp_forward")
        return y

    def loss_grad(self, x: torch.Tensor, target: torch.Tensor, split_weight: float = 1.0):
        x = x.to(self.device, torch.float32).contiguous().reshape(-1, _lib.MLP_IN)
        target = target.to(self.device, torch.float32).contiguous().reshape(-1, _lib.MLP_OUT)
        n = x.shape[0]
        need = lib().fgs_mlp_grad_ws_size(n)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=self.device)
        check(lib().fgs_mlp_loss_grad(_p(self.params), _p(x), _p(target), n, split_weight,
                                      _p(self.grad), _p(self.loss), _p(self._ws), self._ws.numel(),
                                      _stream(self.device)), "fgs_mlp_loss_grad")
        return self.loss, self.grad

    def step(self, x, target, split_weight: float = 1.0) -> torch.Tensor:
        """One Adam step on a batch; returns the batch loss (device tensor)."""
        self.loss_grad(x, target, split_weight)
        self.steps += 1
        check(lib().fgs_adam_update(_p(self.params), _p(self.grad), _p(self.m), _p(self.v),
                                    self.params.numel(), self.lr, 0.9, 0.999, 1e-8, self.steps,
                                    _stream(self.device)), "fgs_adam_update")
        return self.loss


class TileSchedule:
    """What ``RasterOptions.tile_sched`` takes."""

    def __init__(self, mode: int, split_len: int = 0, model: CostModel | None = None,
                 features: bool = False, log_cycles: bool = False, order: bool = False,
                 ahead: bool = False):
        if mode == _lib.TILES_LEARNED and model is None:
            raise ValueError("the learned schedule needs a CostModel")
        if ahead and mode != _lib.TILES_LEARNED:
            raise ValueError("ahead needs the learned schedule")
        self.mode, self.split_len, self.model = mode, int(split_len), model
        self.features, self.log_cycles, self.order = features, log_cycles, order
        # ahead: each frame dispatches the longest-predicted-first work list
        # that fgs_preprocess builds, beside the preprocess, from the tile
        # features the previous frame of the same size computed beside its
        # compositing; no prediction runs on a frame's critical path (the
        # first frame, or one after a size change, runs in tile order).
        self.ahead = ahead
        self._ahead = None  # [T, features (T x 8), work list (2T + 1 words), features valid]

    def list_addresses(self, T: int, device) -> tuple:
        """Device addresses (items, n_items, prev_features, features_out) for
        a frame of T tiles; 0 where there is none."""
        if not self.ahead:
            return (0, 0, 0, 0)
        st = self._ahead
        if st is None or st[0] != T or st[1].device != device:
            st = self._ahead = [T, torch.zeros(T, _lib.MLP_IN, device=device),
                                torch.zeros(2 * T + 1, dtype=torch.int32, device=device), False]
        _, feat, lst, valid = st
        if not valid:
            return (0, 0, 0, feat.data_ptr())
        return (lst.data_ptr(), lst.data_ptr() + 8 * T, feat.data_ptr(), feat.data_ptr())

    def advance(self):
        """After a frame was queued: its features feed the next frame's list."""
        if self.ahead and self._ahead is not None:
            self._ahead[3] = True

    @property
    def uses_list(self) -> bool:
        return self.mode == _lib.TILES_LEARNED and self.order

    @classmethod
    def raster(cls, **kw):
        return cls(_lib.TILES_RASTER, **kw)

    @classmethod
    def heuristic(cls, split_len: int = 4096, **kw):
        return cls(_lib.TILES_HEURISTIC, split_len=split_len, **kw)

    @classmethod
    def learned(cls, model: CostModel, **kw):
        return cls(_lib.TILES_LEARNED, model=model, **kw)

    def to_c(self) -> TileSched:
        """The C struct without work lists (the stage-by-stage paths; the
        ahead lists go through Frame.forward)."""
        return TileSched(self.mode, self.split_len,
                         _p(self.model.params) if self.model is not None else None,
                         1 if self.features else 0, 1 if self.log_cycles else 0,
                         1 if self.order else 0, None, None, None, None)


# Workgroups of the compositing pass resident at once on an MI355X: 256 CUs x
# 6 (its 25 KB of LDS per workgroup allows 6 per CU). A tile whose duration
# exceeds the mean load per slot cannot be balanced without splitting it.
RESIDENT_SLOTS = 256 * 6


def tile_targets(frame, slots: int = RESIDENT_SLOTS):
    """(features T x 8, targets T x 2) of a frame rendered with features and
    log_cycles on: target 0 = log2(1 + ticks of the tile's items), target 1 =
    1 when the tile's ticks exceed the total over ``slots``."""
    v = frame.view_arrays()
    sched = frame.opts.tile_sched
    T = v["tile_features"].shape[0] // 8
    feat = v["tile_features"].reshape(T, 8)
    ticks = v["item_ticks"].to(torch.int64) & 0xFFFFFFFF
    dev = ticks.device
    if sched.uses_list:  # work items
        n_items = int(v["counts"][4].item())
        tile = (v["work_items"][:n_items].to(torch.int64) & 0xFFFFFFFF) >> 2
    elif sched.mode == _lib.TILES_RASTER:  # one workgroup per tile
        n_items = T
        tile = torch.arange(T, device=dev)
    else:  # two grid slots per tile, logged at slot * T + tile
        n_items = 2 * T
        tile = torch.arange(2 * T, device=dev) % T
    cost = torch.zeros(T, dtype=torch.float64, device=ticks.device)
    cost.index_add_(0, tile, ticks[:n_items].to(torch.float64))
    per_slot = cost.sum() / slots
    tgt = torch.stack([torch.log2(1.0 + cost), (cost > per_slot).to(torch.float64)], 1)
    return feat.float(), tgt.float(), cost


class OnlineTrainer:
    """Ring buffer of (tile features, targets) from logged frames, trained
    with the HIP kernels of the cost model."""

    def __init__(self, model: CostModel, capacity: int = 1 << 16, batch: int = 4096,
                 seed: int = 0):
        self.model = model
        dev = model.device
        self.x = torch.zeros(capacity, _lib.MLP_IN, device=dev)
        self.t = torch.zeros(capacity, _lib.MLP_OUT, device=dev)
        self.size = 0
        self.head = 0
        self.batch = batch
        self.gen = torch.Generator(device=dev).manual_seed(seed)

    def observe(self, frame):
        feat, tgt, _ = tile_targets(frame)
        n = feat.shape[0]
        cap = self.x.shape[0]
        idx = (torch.arange(n, device=feat.device) + self.head) % cap
        self.x[idx] = feat
        self.t[idx] = tgt
        self.head = (self.head + n) % cap
        self.size = min(cap, self.size + n)

    def train(self, steps: int) -> float:
        loss = None
        for _ in range(steps):
            idx = torch.randint(0, self.size, (min(self.batch, self.size),),
                                device=self.x.device, generator=self.gen)
            loss = self.model.step(self.x[idx], self.t[idx])
        return float(loss.item()) if loss is not None else float("nan")
