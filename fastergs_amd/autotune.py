"""Autotuner of the FAST schedule's knobs.

The knobs change the work decomposition, never the rendered values:
``bucket_size`` (splats per bucket: compositing parallelism against
checkpoint traffic and backward restarts), ``split_tiles`` (rect size from
which a Gaussian's tile walk is spread over a whole workgroup instead of
being lane-packed) and the tile schedule of the compositing pass. The tile
shape is fixed at 16x16 pixels (as in the reference's render_16x16), so it
is not a knob here.

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

DEFAULT_BUCKETS = (16, 32, 64)
DEFAULT_SPLITS = (16, 32, 64, 128, 256)


def _describe(o: RasterOptions) -> dict:
    d = {"bucket_size": o.bucket_size, "split_tiles": o.split_tiles}
    ts = o.tile_sched
    d["tiles"] = "raster" if ts is None else {0: "raster", 1: f"heuristic({ts.split_len})",
                                              2: "learned"}[ts.mode]
    return d


class Autotuner:
    def __init__(self, buckets=DEFAULT_BUCKETS, splits=DEFAULT_SPLITS, tile_scheds=(None,),
                 reps: int = 3, cache_path: str | None = None):
        self.candidates = [RasterOptions(schedule=_lib.SCHEDULE_FAST, bucket_size=b,
                                         split_tiles=s, tile_sched=t)
                           for b, s, t in itertools.product(buckets, splits, tile_scheds)]
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
        """Returns (best RasterOptions, [(description, ms), ...] sorted)."""
        if key is not None and key in self.cache:
            c = self.cache[key]
            best = RasterOptions(schedule=_lib.SCHEDULE_FAST, bucket_size=c["bucket_size"],
                                 split_tiles=c["split_tiles"])
            return best, c.get("table", [])
        table = []
        for o in self.candidates:
            try:
                table.append((o, self._time(step, o)))
            except torch.cuda.OutOfMemoryError:  # this decomposition does not fit: skip it
                torch.cuda.empty_cache()
        table.sort(key=lambda t: t[1])
        best = table[0][0]
        rows = [(_describe(o), round(ms, 4)) for o, ms in table]
        if key is not None:
            self.cache[key] = dict(_describe(best), table=rows)
            self.cache[key].pop("tiles", None)
            if self.cache_path:
                with open(self.cache_path, "w") as f:
                    json.dump(self.cache, f, indent=1)
        return best, rows
