"""Autotuner of the FAST schedule's knobs.

Knobs that never change the rendered values or gradients (they only change
the work decomposition):

* ``split_tiles``: rect size (in tiles) from which a Gaussian's tile walk is
  spread over row bands of whole waves instead of being lane-packed;
* the compositing work shape: every 16x16 tile as one work item, or tiles
  with at least ``split_len`` list entries as two 16x8 row bands (the tile
  schedule of the compositing pass; ``split_len`` 1 splits every tile);
* the compositing launch: one workgroup per work item, or a persistent grid
  taking the work items of every view from per-XCD work-queue counters
  (``composite_queue``).

Semantic knobs, tuned only on request:

* the bucket size: the cut-off is tested at bucket ends, so last
  contributors, images (by < 1e-4 of a colour) and gradients depend on it;
* the tile shape (``tile_w``: 8 x 32, 16 x 16 or 32 x 8 pixels; the
  reference hard-codes 16 x 16, render_16x16, __init__.py:109): the per-tile
  lists, and with them the bucket boundaries, are defined on the tile.

By default the tuner keeps the bucket size and tile shape of ``base`` (the
reference's 128 and 16 x 16 unless the caller chose others); passing
``buckets`` or ``tile_ws`` makes them tuned knobs too, for callers that
accept that the result then depends on the choice.

``Autotuner.tune(step, key)`` times ``step(options)`` -- one full training
step with those options, e.g. forward + backward of a multi-view batch --
for every candidate with HIP events on the current stream (a warm-up call,
then the median of ``reps`` timed calls), returns the fastest options and
caches the choice under ``key`` (for instance the Gaussian count and the
view sizes), optionally in a JSON file so later runs reuse it.
"""
from __future__ import annotations

import itertools
import json
import os
import statistics

import torch

from . import _lib
from .rasterizer import RasterOptions
from .scheduler import TileSchedule

DEFAULT_SPLITS = (16, 32, 64, 128, 256)
# compositing work shapes: None = whole 16x16 tiles; n = 16x8 row bands for
# tiles with at least n list entries
DEFAULT_WORK_SPLITS = (None, 1)


def _sched(split_len):
    return None if split_len is None else TileSchedule.heuristic(split_len=split_len)


def _describe(o: RasterOptions) -> dict:
    ts = o.tile_sched
    return {"bucket_size": o.bucket_size, "tile_w": int(o.tile_w), "split_tiles": o.split_tiles,
            "band_split_len": None if ts is None else int(ts.split_len),
            "composite_queue": bool(o.composite_queue)}


def _from_description(c: dict) -> RasterOptions:
    return RasterOptions(schedule=_lib.SCHEDULE_FAST, bucket_size=int(c["bucket_size"]),
                         split_tiles=int(c["split_tiles"]),
                         tile_sched=_sched(c.get("band_split_len")),
                         composite_queue=bool(c.get("composite_queue", False)),
                         tile_w=int(c.get("tile_w", 16)))


class Autotuner:
    def __init__(self, splits=DEFAULT_SPLITS, work_splits=DEFAULT_WORK_SPLITS, buckets=None,
                 base: RasterOptions | None = None, reps: int = 3,
                 cache_path: str | None = None, queues=(False, True), tile_ws=None):
        base = base or RasterOptions()
        bucket_list = tuple(buckets) if buckets else (base.bucket_size,)
        tile_list = tuple(tile_ws) if tile_ws else (base.tile_w,)
        self.candidates = [RasterOptions(schedule=_lib.SCHEDULE_FAST, bucket_size=b, tile_w=tw,
                                         split_tiles=s, tile_sched=_sched(w),
                                         composite_queue=q, speculative=base.speculative)
                           for b, tw, s, w, q in itertools.product(bucket_list, tile_list, splits,
                                                                   work_splits, queues)]
        self.reps = reps
        self.cache_path = cache_path
        self.cache: dict = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.cache = json.load(f)

    def _time(self, step, opts) -> float:
        step(opts)  # warm-up (allocations, first-touch)
        times = []
        for _ in range(self.reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            step(opts)
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        return statistics.median(times)

    def tune(self, step, key: str | None = None):
        """Returns (best RasterOptions, [(description, ms), ...] sorted). A
        cached choice is returned with every knob it was tuned with."""
        if key is not None and key in self.cache:
            c = self.cache[key]
            return _from_description(c), c.get("table", [])
        table = []
        for o in self.candidates:
            try:
                table.append((o, self._time(step, o)))
            except torch.cuda.OutOfMemoryError:  # this decomposition does not fit: skip it
                torch.cuda.empty_cache()
        if not table:
            raise RuntimeError("no candidate fitted in memory")
        table.sort(key=lambda t: t[1])
        best = table[0][0]
        rows = [(_describe(o), round(ms, 4)) for o, ms in table]
        if key is not None:
            self.cache[key] = dict(_describe(best), table=rows)
            if self.cache_path:
                with open(self.cache_path, "w") as f:
                    json.dump(self.cache, f, indent=1)
        return best, rows
