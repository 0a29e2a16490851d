"""Drop-in Python API of the FasterGS rasterizer on gfx950.

Mirrors gs_submodules/diff-gaussian-rasterization_faster/
diff_gaussian_rasterization_faster/__init__.py of the FasterGS repository:
``GaussianRasterizationSettings`` (:191-205), ``GaussianRasterizer`` (:207-264),
``rasterize_gaussians`` (:41-73) and the autograd function (:76-187) with the
same argument order, defaults, return values and in-place densification
statistics. Differences, all deliberate:

* workspaces are per call and kept by the autograd context (the reference
  shares one static 2.4 GB buffer between calls, __init__.py:21-38, :77, so a
  second forward before a backward corrupts the first);
* there is no pair cap: buffers are sized from the exact pair count (the
  reference raises a ``str`` above 2^24 pairs, __init__.py:118-119);
* ``cov3D_precomp`` is rejected with ``ValueError`` (the reference accepts it
  and ignores it, csrc/pybind.cpp:26-35);
* argument errors raise ``ValueError`` instead of a bare ``Exception``.

Compute runs in libfastergs.so (hand-written gfx950 kernels) through its C
ABI: a frame's forward and backward go through the PyTorch-ROCm extension
``fastergs_amd._C`` (torch-typed ops over the ABI, the role of the
reference's ``_C.ops``); the stage-level and inspection calls of ``Frame``
use the ABI directly through ctypes. PyTorch only provides device memory
and the current stream.
"""
from __future__ import annotations

import collections
import ctypes
from dataclasses import dataclass
from typing import NamedTuple

import torch
import torch.nn as nn

from . import _C, _lib
from ._lib import Camera, FrameView, Gaussians, Grads, Options, check, lib

_NONE = torch.empty(0)  # an absent optional tensor for the extension's ops


@dataclass
class RasterOptions:
    """Schedule knobs (not part of the reference settings tuple).

    The defaults are the reference backend's semantics: the bucketed FAST
    schedule with 128 splats per bucket (BUCKET_SIZE, rasterizer_imp.h:24),
    so an unchanged caller gets FasterGS's images, cut-offs and gradients.
    ``bucket_size`` is semantic: the T < 1e-4 cut-off is tested at bucket
    ends, so last contributors and gradients depend on it.
    ``tile_sched``: work schedule of the FAST compositing pass, an object
    with a ``to_c()`` returning a ``_lib.TileSched`` (see
    ``fastergs_amd.scheduler``); None = one workgroup per tile in tile order.
    """
    schedule: int = _lib.SCHEDULE_FAST
    bucket_size: int = 128
    split_tiles: int = 64
    tile_sched: object = None
    # FAST compositing as a persistent grid taking work items from a device
    # counter (same outputs as one workgroup per item)
    composite_queue: bool = False
    # tile width in pixels (tiles of 256 pixels: 8 x 32, 16 x 16 or 32 x 8);
    # 16 is the reference's 16 x 16. Semantic in FAST like bucket_size: the
    # per-tile lists and bucket boundaries are defined on the tile
    tile_w: int = 16
    # FAST single views through the drop-in API: build the per-tile lists of
    # the nearest ceil(P / div) Gaussians first, starting at div =
    # list_prefix_div (0 or 1: never). Every pixel that stops at the end of a
    # whole bucket within them has its final values; if some pixel does not,
    # the whole lists are built and the frame composited again on the device.
    # Outputs never depend on it. Per image shape the divisor then follows
    # the outcomes (see _PrefixPolicy): larger while prefixes suffice, smaller
    # when they do not.
    list_prefix_div: int = 16
    # render with pair / bucket capacities remembered from earlier frames of
    # the same shape instead of waiting for this frame's counts (the counts
    # are checked while the render runs; an overflow renders again exactly)
    speculative: bool = True

    def to_c(self, abs_grad: bool = True) -> Options:
        o = Options(self.schedule, self.bucket_size, self.split_tiles, 1 if abs_grad else 0, None)
        o.composite_queue = 1 if self.composite_queue else 0
        o.tile_w = int(self.tile_w)
        if self.tile_sched is not None:
            ts = self.tile_sched.to_c()
            o.tile_sched = ctypes.pointer(ts)
            o._keep = ts  # the struct must outlive the calls that read it
        return o


_default_options = RasterOptions()


def set_default_options(**kw) -> RasterOptions:
    for k, v in kw.items():
        if not hasattr(_default_options, k):
            raise ValueError(f"unknown option {k}")
        setattr(_default_options, k, v)
    return _default_options


def get_default_options() -> RasterOptions:
    return _default_options


class GaussianRasterizationSettings(NamedTuple):
    image_height: int
    image_width: int
    tanfovx: float
    tanfovy: float
    bg: torch.Tensor
    scale_modifier: float
    viewmatrix: torch.Tensor
    projmatrix: torch.Tensor
    sh_degree: int
    campos: torch.Tensor
    mult: float
    prefiltered: bool
    debug: bool


def _ptr(t: torch.Tensor | None):
    return None if t is None or t.numel() == 0 else ctypes.c_void_p(t.data_ptr())


def _f32(t: torch.Tensor, device) -> torch.Tensor:
    return t.to(device=device, dtype=torch.float32).contiguous()


def _stream(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


# Largest pair count / bucket bound seen per frame shape (plus headroom):
# the capacity a speculative forward renders with.
_capacity_hints: dict = {}


class _PrefixPolicy:
    """Prefix size of the frames of one shape, adapted from their outcomes.

    Each prefix frame's status word (pinned host memory that compositing
    sets to 1 on the device when the prefix lists did not suffice) is read
    once an event after the frame has completed, so the host never waits.
    After STREAK frames that sufficed, one frame tries a divisor twice as
    large (a prefix half as long); the frames in between keep the last one
    that sufficed. A tried divisor that did not suffice caps later tries
    below it; a frame at the kept divisor that did not suffice halves it.
    Below divisor 2, or when no prefix of the shape has sufficed yet, frames
    render whole lists for RETRY frames before starting over."""
    STREAK = 4
    MAX_DIV = 256
    RETRY = 256

    def __init__(self, div: int):
        self.start = self.good = div
        self.cap = self.MAX_DIV
        self.streak = 0
        self.probe = 0      # divisor of the try in flight (0: none)
        self.off = 0        # frames left with whole lists
        self.ever_ok = False
        self.pending = collections.deque()  # (status word, event, divisor) in frame order
        self.free = []

    def _outcome(self, div: int, failed: bool):
        self.ever_ok |= not failed
        if div == self.probe:
            self.probe = 0
            if failed:
                self.cap = div // 2
            else:
                self.good, self.streak = div, 0
        elif failed:
            if div == self.good:
                self.good //= 2
                self.cap = max(self.good, 1)
                self.streak = 0
                if self.good < 2 or not self.ever_ok:
                    self.off = self.RETRY
        elif div == self.good:
            self.streak += 1

    def divisor(self) -> int:
        """Divisor for the next frame (0: whole lists)."""
        while self.pending and self.pending[0][1].query():
            status, _, div = self.pending.popleft()
            failed = int(status[0]) != 0
            status.zero_()
            self.free.append(status)
            self._outcome(div, failed)
        if self.off > 0:
            self.off -= 1
            if self.off == 0:
                self.good, self.cap, self.streak = self.start, self.MAX_DIV, 0
                self.ever_ok = False
            return 0
        if not self.probe and self.streak >= self.STREAK and 2 * self.good <= self.cap:
            self.probe = 2 * self.good
            return self.probe
        return self.good

    def status_word(self) -> torch.Tensor:
        return self.free.pop() if self.free else torch.zeros(1, dtype=torch.int32,
                                                             pin_memory=True)

    def track(self, status: torch.Tensor, div: int):
        ev = torch.cuda.Event()
        ev.record()
        self.pending.append((status, ev, div))


_prefix_policies: dict = {}


class Frame:
    """One forward evaluation: inputs, workspaces and outputs.

    Owned by the autograd context between forward and backward; tests use it
    to read intermediate products (sorted lists, ranges, buckets).
    """

    def __init__(self, settings: GaussianRasterizationSettings, means3D, opacities, scales,
                 rotations, dc=None, shs=None, colors_precomp=None, culling=None,
                 options: RasterOptions | None = None, abs_grad: bool = True,
                 forward_only: bool = False):
        if not means3D.is_cuda:
            raise ValueError("means3D must be a GPU tensor")
        dev = means3D.device
        self.device = dev
        self.opts = options or _default_options
        self.copts = self.opts.to_c(abs_grad)
        # a frame that will not be differentiated keeps no bucket checkpoints
        self.forward_only = bool(forward_only)
        self.copts.forward_only = 1 if forward_only else 0
        self.P = P = int(means3D.shape[0])
        self.W, self.H = int(settings.image_width), int(settings.image_height)
        self.settings = settings
        self.means3D = _f32(means3D, dev).reshape(P, 3)
        self.opacities = _f32(opacities, dev).reshape(P)
        self.scales = _f32(scales, dev).reshape(P, 3)
        self.rotations = _f32(rotations, dev).reshape(P, 4)
        self.colors = None
        self.dc = self.shs = None
        self.M = 0
        if colors_precomp is not None and colors_precomp.numel() > 0:
            self.colors = _f32(colors_precomp, dev).reshape(P, 3)
        else:
            if P > 0 and (dc is None or dc.numel() == 0):
                raise ValueError("dc is required when colours come from SH")
            self.dc = _f32(dc, dev).reshape(P, 3) if P > 0 else None
            if shs is not None and shs.numel() > 0:
                self.shs = _f32(shs, dev).reshape(P, -1, 3)
                self.M = int(self.shs.shape[1])
        self.culling = None
        if culling is not None and culling.numel() > 0:
            self.culling = culling.to(device=dev, dtype=torch.bool).contiguous().reshape(P)
        self.view = _f32(settings.viewmatrix, dev).reshape(16)
        self.proj = _f32(settings.projmatrix, dev).reshape(16)
        self.campos = _f32(settings.campos, dev).reshape(3)
        bg = settings.bg if torch.is_tensor(settings.bg) else torch.tensor(settings.bg)
        self.bg = _f32(bg, dev).reshape(3)
        self.R = self.NB = 0          # this frame's pair count and bucket-table bound
        self.R_cap = self.NB_cap = 0  # capacities the pair workspace is laid out for
        self.pws = None

    # ctypes views of the frame for the stage-level ABI calls
    @property
    def cam(self) -> Camera:
        st = self.settings
        c = Camera(self.W, self.H, float(st.tanfovx), float(st.tanfovy), float(st.scale_modifier),
                   float(st.mult), int(st.sh_degree), _ptr(self.view), _ptr(self.proj),
                   _ptr(self.campos), _ptr(self.bg))
        return c

    @property
    def gauss(self) -> Gaussians:
        return Gaussians(self.P, self.M, _ptr(self.means3D), _ptr(self.opacities),
                         _ptr(self.scales), _ptr(self.rotations), _ptr(self.dc), _ptr(self.shs),
                         _ptr(self.colors), _ptr(self.culling))

    def _ext_args(self, lists: bool = False):
        """The frame's tensors, settings and options as the extension's ops
        take them; `lists`: with the tile schedule's work-list addresses (the
        forward of an ahead schedule)."""
        st, o = self.settings, self.copts
        ts = self.opts.tile_sched
        mlp = _NONE
        if ts is None:
            tsl = [-1, 0, 0, 0, 0, 0, 0, 0, 0]
        else:
            T = _lib.tiles_of(self.W, self.H, self.opts.tile_w)
            addr = (ts.list_addresses(T, self.device)
                    if lists and hasattr(ts, "list_addresses") else (0, 0, 0, 0))
            tsl = [int(ts.mode), int(ts.split_len), int(bool(ts.features)),
                   int(bool(ts.log_cycles)), int(bool(ts.order)), *(int(a) for a in addr)]
            if getattr(ts, "model", None) is not None:
                mlp = ts.model.params
        status = getattr(self, "list_status", None)
        opts = [int(o.schedule), int(o.bucket_size), int(o.split_tiles), int(o.abs_grad),
                int(o.forward_only), *tsl, int(o.composite_queue), int(o.tile_w),
                int(getattr(self, "list_prefix", 0)), 0 if status is None else status.data_ptr()]
        n = lambda t: _NONE if t is None else t
        return (self.means3D, self.opacities, self.scales, self.rotations, n(self.dc), n(self.shs),
                n(self.colors), n(self.culling), self.view, self.proj, self.campos, self.bg,
                self.W, self.H, float(st.tanfovx), float(st.tanfovy), float(st.scale_modifier),
                float(st.mult), int(st.sh_degree), opts, mlp)

    # -- forward stages ---------------------------------------------------
    def preprocess(self, sync: bool = True):
        L, dev = lib(), self.device
        self.radii = torch.empty(self.P, dtype=torch.int32, device=dev)  # every entry written
        n = L.fgs_gauss_ws_size(self.P, self.W, self.H, ctypes.byref(self.copts))
        self.gws = torch.empty(n, dtype=torch.uint8, device=dev)
        check(L.fgs_preprocess(ctypes.byref(self.cam), ctypes.byref(self.gauss),
                               ctypes.byref(self.copts), _ptr(self.gws), n, _ptr(self.radii),
                               ctypes.c_void_p(_stream(dev))), "fgs_preprocess")
        if sync:
            self.read_counts(torch.cuda.current_stream(dev))
            self.R_cap, self.NB_cap = self.R, self.NB
        return self

    def read_counts(self, stream):
        """Pair count R and bucket-table capacity NB of the preprocessed frame
        (waits on `stream` only)."""
        counts = (ctypes.c_int64 * 2)()
        check(lib().fgs_read_counts(_ptr(self.gws), self.P, self.W, self.H,
                                    ctypes.byref(self.copts), counts,
                                    ctypes.c_void_p(stream.cuda_stream)), "fgs_read_counts")
        self.R, self.NB = int(counts[0]), int(counts[1])

    def render(self):
        L, dev = lib(), self.device
        n = L.fgs_pair_ws_size(self.P, self.W, self.H, self.R_cap, self.NB_cap,
                               ctypes.byref(self.copts))
        self.pws = torch.empty(n, dtype=torch.uint8, device=dev)
        self.out_color = torch.empty((3, self.H, self.W), dtype=torch.float32, device=dev)
        self.T_final = torch.empty((self.H, self.W, 1), dtype=torch.float32, device=dev)
        self.last_contributor = torch.empty((self.H, self.W, 1), dtype=torch.int32, device=dev)
        check(L.fgs_render(ctypes.byref(self.cam), ctypes.byref(self.gauss),
                           ctypes.byref(self.copts), _ptr(self.gws), self.gws.numel(),
                           _ptr(self.pws), n, self.R_cap, self.NB_cap, _ptr(self.out_color),
                           _ptr(self.T_final), _ptr(self.last_contributor),
                           ctypes.c_void_p(_stream(dev))), "fgs_render")
        return self

    def forward(self):
        """Preprocess and render (one call of the extension). With speculative
        options and a capacity remembered for frames of this shape, the render
        is queued at once with that capacity while the frame's counts are
        read on a side stream that waits for the preprocess only; a frame that
        exceeds the capacity is rendered again with its exact counts (the
        outputs never depend on the capacity)."""
        key = (self.P, self.W, self.H, self.opts.schedule, self.opts.bucket_size, self.device)
        hint = _capacity_hints.get(key) if self.opts.speculative else None
        R_hint, NB_hint = hint if hint is not None else (0, 0)
        pol, div = None, int(self.opts.list_prefix_div)
        if div > 1 and self.opts.schedule == _lib.SCHEDULE_FAST and self.opts.tile_sched is None:
            # per image shape, not per Gaussian count: the prefix is a share
            # of the Gaussians, and its size carries over when densification
            # changes their number
            pkey = (self.W, self.H, self.opts.bucket_size, self.opts.tile_w, self.device, div)
            pol = _prefix_policies.setdefault(pkey, _PrefixPolicy(div))
            div = pol.divisor()
            if div > 1 and self.P > div:
                self.list_prefix = -(-self.P // div)
                self.list_status = pol.status_word()
            else:
                pol = None
        (self.radii, self.gws, self.pws, self.out_color, self.T_final, self.last_contributor,
         c, self.grad_ws) = _C.ops.forward(*self._ext_args(lists=True), int(R_hint), int(NB_hint))
        if pol is not None:
            pol.track(self.list_status, div)
            self.list_status = None  # later calls on the frame do not report
            self.prefix_lists = True
        ts = self.opts.tile_sched
        if ts is not None and hasattr(ts, "advance"):
            ts.advance()
        self.R, self.NB, self.R_cap, self.NB_cap = (int(x) for x in c)
        # remember the largest counts seen for this shape, with headroom
        old = _capacity_hints.get(key, (0, 0))
        _capacity_hints[key] = (max(old[0], self.R + self.R // 8 + 1024),
                                max(old[1], self.NB + self.NB // 8 + 64))
        return self

    def list_counts(self) -> dict:
        """Pair counts of the frame's list build (waits for the frame):
        ``pairs`` of the whole lists, ``prefix_pairs`` of the prefix lists
        (0 when the frame was built with whole lists) and ``redo`` whether
        the prefixes did not suffice and the whole lists were built too."""
        v = FrameView()
        check(lib().fgs_frame_view_get(_ptr(self.gws), None, self.P, self.W, self.H, 0, 0,
                                       ctypes.byref(self.copts), ctypes.byref(v)),
              "fgs_frame_view_get")
        c = _tensor_at(v.counts, 16, torch.int32, self).cpu().numpy().view("uint32")
        return {"pairs": int(c[0]), "prefix_pairs": int(c[7]), "redo": bool(c[8])}

    def complete_lists(self):
        """Whole per-tile lists for a frame rendered from prefix lists (the
        list passes and inspection views read whole lists): emission, sort,
        ranges and compositing again with whole lists into the frame's own
        workspaces and outputs, which keep their values."""
        if not getattr(self, "prefix_lists", False):
            return self
        self.list_prefix = 0
        check(lib().fgs_render(ctypes.byref(self.cam), ctypes.byref(self.gauss),
                               ctypes.byref(self.copts), _ptr(self.gws), self.gws.numel(),
                               _ptr(self.pws), self.pws.numel(), self.R_cap, self.NB_cap,
                               _ptr(self.out_color), _ptr(self.T_final),
                               _ptr(self.last_contributor), ctypes.c_void_p(_stream(self.device))),
              "fgs_render")
        self.prefix_lists = False
        return self

    # -- backward -----------------------------------------------------------
    def backward(self, grad_out_color: torch.Tensor, skip=(), stats=None) -> dict:
        """Gradients of the frame; `skip` may name "conic", "colors" and
        "cov3D", which are then neither computed into memory nor returned.
        `stats`: (xyz_gradient_accum, xyz_gradient_accum_abs, max_radii,
        denom) (entries may be None) updated in place as densify_stats
        would, in the same pass."""
        if not set(skip) <= {"conic", "colors", "cov3D"}:
            raise ValueError("only conic, colors and cov3D gradients can be skipped")
        if self.forward_only:
            raise ValueError("the frame was rendered forward-only")
        # the workspace the forward's preprocess zeroed serves one backward
        zeroed, self.grad_ws = getattr(self, "grad_ws", None), None
        r = _C.ops.backward(*self._ext_args(), self.gws, self.pws, self.R_cap, self.NB_cap,
                            self.radii, self.T_final, self.last_contributor, grad_out_color,
                            "conic" not in skip, "colors" not in skip, "cov3D" not in skip,
                            [_NONE if t is None else t for t in (stats or (None,) * 4)],
                            _NONE if zeroed is None else zeroed)
        names = ("mean2D", "conic", "opacities", "colors", "means3D", "cov3D", "dc", "shs",
                 "scales", "rotations")
        out = {k: t for k, t in zip(names, r[:10]) if k not in skip}
        return _Grads(out, r[10], self.P)

    def densify_stats(self, dmean2D, accum, accum_abs, max_radii, denom):
        n = lambda t: _NONE if t is None else t
        _C.ops.densify_stats(self.radii, dmean2D, n(accum), n(accum_abs), n(max_radii), n(denom))

    def render_simp(self):
        """Importance statistics over this frame's lists (render_simp):
        (accum_weights (P,) float, accum_count (P,) int32, accum_max_count (P,) float)."""
        self.complete_lists()
        dev, P = self.device, self.P
        w = torch.empty(P, dtype=torch.float32, device=dev)
        c = torch.empty(P, dtype=torch.int32, device=dev)
        m = torch.empty(P, dtype=torch.float32, device=dev)
        check(lib().fgs_render_simp(_ptr(self.gws), _ptr(self.pws), P, self.W, self.H, self.R_cap,
                                    self.NB_cap, ctypes.byref(self.copts), _ptr(w), _ptr(c), _ptr(m),
                                    ctypes.c_void_p(_stream(dev))), "fgs_render_simp")
        return w, c, m

    def render_depth(self) -> dict:
        """Per-pixel depth over this frame's lists (render_depth): out_pts
        (3,H,W), rendered_depth (1,H,W), accum_alpha (1,H,W), gidx (1,H,W)
        int32 and discriminants (1,H,W)."""
        self.complete_lists()
        dev, H, W = self.device, self.H, self.W
        e = lambda c, dt=torch.float32: torch.empty((c, H, W), dtype=dt, device=dev)
        out = {"out_pts": e(3), "rendered_depth": e(1), "accum_alpha": e(1),
               "gidx": e(1, torch.int32), "discriminants": e(1)}
        check(lib().fgs_render_depth(_ptr(self.gws), _ptr(self.pws), self.P, W, H, self.R_cap,
                                     self.NB_cap, ctypes.byref(self.copts), _ptr(self.means3D),
                                     _ptr(self.scales), _ptr(self.rotations), _ptr(self.proj),
                                     _ptr(self.campos), _ptr(out["out_pts"]),
                                     _ptr(out["rendered_depth"]), _ptr(out["accum_alpha"]),
                                     _ptr(out["gidx"]), _ptr(out["discriminants"]),
                                     ctypes.c_void_p(_stream(dev))), "fgs_render_depth")
        return out

    # -- inspection -----------------------------------------------------------
    def view_arrays(self) -> dict:
        """Torch views of the intermediate arrays (valid until the frame dies)."""
        self.complete_lists()
        if self.opts.schedule == _lib.SCHEDULE_FAST and getattr(self, "pws", None) is not None:
            # the bucket tables are built on demand (no pass of the step reads them)
            check(lib().fgs_bucket_tables(_ptr(self.gws), _ptr(self.pws), self.P, self.W, self.H,
                                          self.R_cap, self.NB_cap, ctypes.byref(self.copts),
                                          ctypes.c_void_p(_stream(self.device))),
                  "fgs_bucket_tables")
        v = FrameView()
        check(lib().fgs_frame_view_get(_ptr(self.gws), _ptr(getattr(self, "pws", None)), self.P,
                                       self.W, self.H, self.R_cap, self.NB_cap,
                                       ctypes.byref(self.copts),
                                       ctypes.byref(v)), "fgs_frame_view_get")
        T = _lib.tiles_of(self.W, self.H, self.opts.tile_w)
        spec = {
            "geom": (torch.float32, self.P * 8), "splat": (torch.float32, self.P * 12),
            "pair_count": (torch.int32, self.P), "tile_count": (torch.int32, T),
            "ranges": (torch.int32, 2 * T), "bucket_off": (torch.int32, T),
            "max_contrib": (torch.int32, T), "counts": (torch.int32, 16),
            "sorted_keys": (torch.int64, self.R), "sorted_vals": (torch.int32, self.R),
            "bucket_ranges": (torch.int32, 2 * self.NB), "bucket_tile": (torch.int32, self.NB),
            "acc_t": (torch.float32, self.NB * 256), "acc_c": (torch.float32, self.NB * 256 * 4),
            "pixel_colors": (torch.float32, self.H * self.W * 4),
            "tile_features": (torch.float32, T * 8), "tile_pred": (torch.float32, T * 2),
            "work_items": (torch.int32, 2 * T), "item_ticks": (torch.int32, 2 * T),
        }
        res = {}
        for name, (dt, n) in spec.items():
            addr = getattr(v, name)
            if not addr:
                continue
            res[name] = _tensor_at(addr, n, dt, self)
        return res


class _Grads(dict):
    """Gradient tensors by name; "raw" (the raw per-record sums, see
    raw_sums) is assembled from the gradient workspace on first access."""

    def __init__(self, grads: dict, grad_ws: torch.Tensor, n: int, shape=None):
        super().__init__(grads)
        self._ws, self._n, self._shape = grad_ws, n, shape

    def __missing__(self, key):
        if key != "raw":
            raise KeyError(key)
        raw = raw_sums(self._ws, self._n)
        self["raw"] = raw if self._shape is None else raw.reshape(self._shape)
        return self["raw"]


def raw_sums(grad_ws: torch.Tensor, n: int) -> torch.Tensor:
    """The backward's raw per-record sums (n x 12 float64, fgs_math.h RawField
    order: colour 3, mean2D 2, conic 3, power 1, |mean2D| 2, one zero) from
    the gradient workspace's conic rows (4 doubles) and float rows of the
    other eight fields."""
    con = grad_ws[:n * 32].view(torch.float64).reshape(n, 4)
    lin = grad_ws[n * 32:n * 64].view(torch.float32).reshape(n, 8).double()
    return torch.cat([lin[:, :5], con[:, :3], lin[:, 5:], torch.zeros_like(con[:, :1])], 1)


def _tensor_at(addr: int, n: int, dtype, frame: Frame) -> torch.Tensor:
    """Copy n elements at a device address inside one of the frame's workspaces."""
    item = torch.tensor([], dtype=dtype).element_size()
    for buf in (frame.gws, getattr(frame, "pws", None)):
        if buf is None:
            continue
        base = buf.data_ptr()
        if base <= addr < base + buf.numel():
            off = addr - base
            return buf[off:off + n * item].view(dtype).clone()
    raise ValueError("address outside the frame workspaces")


def rasterize_gaussians(means3D, dc, shs, colors_precomp, opacities, scales, rotations,
                        cov3Ds_precomp, raster_settings, culling, xyz_gradient_accum,
                        xyz_gradient_accum_abs, max_radii, denom):
    # grad mode as the caller has it (autograd runs forward() with it off)
    return _RasterizeGaussians.apply(means3D, dc, shs, colors_precomp, opacities, scales,
                                     rotations, cov3Ds_precomp, raster_settings, culling,
                                     xyz_gradient_accum, xyz_gradient_accum_abs, max_radii, denom,
                                     torch.is_grad_enabled())


class _RasterizeGaussians(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means3D, dc, shs, colors_precomp, opacities, scales, rotations,
                cov3Ds_precomp, raster_settings, culling, xyz_gradient_accum,
                xyz_gradient_accum_abs, max_radii, denom, grad_enabled=True):
        if cov3Ds_precomp is not None and cov3Ds_precomp.numel() > 0:
            raise ValueError("cov3D_precomp is not supported: pass scales and rotations")
        P = int(means3D.shape[0])
        if scales is None or rotations is None or (P > 0 and (scales.numel() == 0 or
                                                             rotations.numel() == 0)):
            raise ValueError("scales and rotations are required")
        want_abs = xyz_gradient_accum_abs is not None and xyz_gradient_accum_abs.numel() > 0
        # called under torch.no_grad, or no input asks for a gradient: nothing
        # to keep for a backward (needs_input_grad reports requires_grad even
        # under no_grad)
        fwd_only = not grad_enabled or not any(ctx.needs_input_grad[:8])
        frame = Frame(raster_settings, means3D, opacities, scales, rotations, dc=dc, shs=shs,
                      colors_precomp=colors_precomp, culling=culling, abs_grad=want_abs,
                      forward_only=fwd_only).forward()
        ctx.frame = frame
        ctx.shapes = [t.shape if torch.is_tensor(t) else None for t in
                      (means3D, dc, shs, colors_precomp, opacities, scales, rotations,
                       cov3Ds_precomp)]
        ctx.stats = (xyz_gradient_accum, xyz_gradient_accum_abs, max_radii, denom)
        ctx.mark_non_differentiable(frame.radii)
        # no zero-filled gradient for the radii (and none for an unused image)
        ctx.set_materialize_grads(False)
        return frame.out_color, frame.radii

    @staticmethod
    def backward(ctx, grad_out_color, _grad_radii):
        frame: Frame = ctx.frame
        if grad_out_color is None:  # the image took no part in the loss
            ctx.frame = None
            return (None,) * 15
        s = ctx.shapes
        # the conic and cov3D gradients are never returned (cov3D_precomp is
        # rejected), the colour gradient only for precomputed colours
        skip = ["conic", "cov3D"]
        if s[3] is None or s[3].numel() == 0:
            skip.append("colors")
        accum, accum_abs, max_radii, denom = ctx.stats
        # the densification statistics (pybind.cpp:324) in the adjoint's pass
        stats = (accum, accum_abs, max_radii, denom) if accum is not None and \
            accum.numel() != 0 else None
        g = frame.backward(grad_out_color, skip=skip, stats=stats)
        ctx.frame = None

        def shaped(t, shape):
            if shape is None or t is None:  # input None (not a tensor), or not computed
                return None
            if t.numel() == shape.numel() and t.numel() > 0:
                return t.reshape(shape)
            return t

        return (shaped(g["means3D"], s[0]), shaped(g["dc"], s[1]), shaped(g["shs"], s[2]),
                shaped(g.get("colors"), s[3]), shaped(g["opacities"], s[4]),
                shaped(g["scales"], s[5]), shaped(g["rotations"], s[6]),
                None, None, None, None, None, None, None, None)


class GaussianRasterizer(nn.Module):
    def __init__(self, raster_settings, **kwargs):
        super().__init__()
        self.raster_settings = raster_settings

    def forward(self, means3D, opacities, dc=None, shs=None, colors_precomp=None, scales=None,
                rotations=None, cov3D_precomp=None, culling=None, xyz_gradient_accum=None,
                xyz_gradient_accum_abs=None, max_radii=None, denom=None):
        if (shs is None and colors_precomp is None) or (shs is not None and colors_precomp is not None):
            raise ValueError("Please provide excatly one of either SHs or precomputed colors!")
        if ((scales is None or rotations is None) and cov3D_precomp is None) or \
                ((scales is not None or rotations is not None) and cov3D_precomp is not None):
            raise ValueError("Please provide exactly one of either scale/rotation pair or "
                             "precomputed 3D covariance!")
        empty = lambda dtype=torch.float32: torch.tensor([], dtype=dtype)
        return rasterize_gaussians(
            means3D, dc if dc is not None else empty(), shs if shs is not None else empty(),
            colors_precomp if colors_precomp is not None else empty(), opacities,
            scales if scales is not None else empty(),
            rotations if rotations is not None else empty(),
            cov3D_precomp if cov3D_precomp is not None else empty(), self.raster_settings,
            culling if culling is not None else empty(torch.bool),
            xyz_gradient_accum if xyz_gradient_accum is not None else empty(),
            xyz_gradient_accum_abs if xyz_gradient_accum_abs is not None else empty(),
            max_radii if max_radii is not None else empty(),
            denom if denom is not None else empty())

    @torch.no_grad()
    def render_simp(self, means3D, means2D, opacities, culling=None, dc=None, shs=None,
                    colors_precomp=None, scales=None, rotations=None, cov3D_precomp=None):
        """Forward render plus the render_simp importance statistics, with the
        argument order of the Mini-Splatting2 backends' ``render_simp``
        (diff_gaussian_rasterization_ms_2/__init__.py:416-455; the FasterGS
        backend binds the same operation as _C.ops.render_simp, pybind.cpp:180-208).
        ``means2D`` is accepted for that signature and unused. Returns
        (color, radii, accum_weights, accum_count, accum_max_count)."""
        if (shs is None and colors_precomp is None) or (shs is not None and colors_precomp is not None):
            raise ValueError("Please provide excatly one of either SHs or precomputed colors!")
        if cov3D_precomp is not None and cov3D_precomp.numel() > 0:
            raise ValueError("cov3D_precomp is not supported: pass scales and rotations")
        if scales is None or rotations is None:
            raise ValueError("scales and rotations are required")
        fr = Frame(self.raster_settings, means3D, opacities, scales, rotations, dc=dc, shs=shs,
                   colors_precomp=colors_precomp, culling=culling, forward_only=True).forward()
        w, c, m = fr.render_simp()
        return fr.out_color, fr.radii, w, c, m

    @torch.no_grad()
    def render_depth(self, means3D, means2D, opacities, culling=None, dc=None, shs=None,
                     colors_precomp=None, scales=None, rotations=None, cov3D_precomp=None):
        """Forward render plus the per-pixel depth of render_depth, with the
        argument order and result dict of the Mini-Splatting2 backends'
        ``render_depth`` (diff_gaussian_rasterization_ms_2/__init__.py:457-497,
        :223-265; the FasterGS backend binds the same kernel as
        _C.ops.render_depth, pybind.cpp:134-172). ``means2D`` is accepted for
        that signature and unused. Returns {"render", "out_pts",
        "rendered_depth", "discriminants", "gidx", "accum_alpha", "radii"}."""
        if (shs is None and colors_precomp is None) or (shs is not None and colors_precomp is not None):
            raise ValueError("Please provide excatly one of either SHs or precomputed colors!")
        if cov3D_precomp is not None and cov3D_precomp.numel() > 0:
            raise ValueError("cov3D_precomp is not supported: pass scales and rotations")
        if scales is None or rotations is None:
            raise ValueError("scales and rotations are required")
        fr = Frame(self.raster_settings, means3D, opacities, scales, rotations, dc=dc, shs=shs,
                   colors_precomp=colors_precomp, culling=culling, forward_only=True).forward()
        res = fr.render_depth()
        res["render"] = fr.out_color
        res["radii"] = fr.radii
        return res
