"""Forward rendering, and whole training steps, captured into HIP graphs.

The render-only loop of the FasterGS callers (3DGS/render_faster.py:35-57)
renders a fixed set of Gaussians from many cameras. Here preprocess, pair
emission, sort, ranges, bucket tables and compositing of one frame are
captured once into a HIP graph (through ``torch.cuda.CUDAGraph``, which is a
hipGraph on ROCm) and replayed per camera:

* the library reads the view / projection matrices and the camera centre
  through device pointers at run time, so a replay renders whatever camera
  was last copied into this object's camera buffers;
* the pair and bucket counts stay on the device: the pair workspace is sized
  once for a capacity (the first frame's counts times ``headroom``) and the
  kernels read the actual counts themselves, so nothing in a frame waits for
  the host;
* every frame folds its counts into a running maximum (inside the graph), so
  ``check()`` can tell after any number of replays whether some frame needed
  more than the capacity -- pairs beyond it are dropped, never written out of
  bounds -- and ``render(..., grow=True)`` re-captures with a larger one.

The FAST and PLAIN schedules are both capturable; outputs of a replay are
bit-identical to an eager ``Frame`` of the same camera.

``CapturedTrainingStep`` captures a whole training step of a batch of views
(one or more, uneven sizes): forward at fixed capacities, the L2 loss
gradient against target buffers, the rasterizer backward and the
densification statistics, replayed after the optimizer has updated the
Gaussians in place.
"""
from __future__ import annotations

import ctypes
import math

import torch

from . import _lib
from ._lib import Camera, FrameView, Gaussians, check, lib
from .rasterizer import GaussianRasterizationSettings, RasterOptions, _f32, _ptr


class CapacityError(RuntimeError):
    """A replayed frame had more pairs or buckets than the captured capacity."""


def _in_place(name: str, t: torch.Tensor, dev, per: int) -> torch.Tensor:
    """A Gaussian tensor that every replay reads in place: it must already be
    contiguous float32 on the capture's device with `per` values per
    Gaussian (a converted copy would not see the caller's later updates)."""
    if not torch.is_tensor(t) or t.dtype != torch.float32 or not t.is_contiguous() or \
            t.device != dev:
        raise ValueError(f"{name} must be a contiguous float32 tensor on {dev} "
                         "(replays read it in place)")
    if per and t.numel() % per:
        raise ValueError(f"{name} has {t.numel()} values, not a multiple of {per}")
    return t


def _prefix_lists(copts, opts: RasterOptions, P: int, V: int, dev):
    """Prefix lists in a captured frame (FAST): the lists of the nearest
    ceil(P V / list_prefix_div) records of the V views first, a size fixed at
    capture (whole lists when the capture's warm-up frame needed them). A replay whose prefixes did not suffice builds the whole lists
    and composites again on the device (same outputs) and sets the returned
    device word to 1 (read and cleared by ``prefix_redone()``)."""
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    div = int(opts.list_prefix_div)
    if opts.schedule == _lib.SCHEDULE_FAST and opts.tile_sched is None and div > 1 and P * V > div:
        copts.list_prefix = -(-(P * V) // div)
        copts.list_status = status.data_ptr()
    return status


def _redone(status: torch.Tensor) -> bool:
    r = bool(int(status.item()))
    status.zero_()
    return r


class _PrefixCapture:
    def _prefix_check(self):
        """After the eager warm-up frame of a capture: a frame whose prefix
        lists did not suffice is captured with whole lists instead (every
        replay would otherwise build both)."""
        if self.copts.list_prefix and _redone(self.prefix_status):
            self.copts.list_prefix = 0


class CapturedRenderer(_PrefixCapture):
    """Forward renderer whose whole frame is one graph replay.

    ``settings`` fixes the image size, field of view, background, SH degree
    and scale modifier; its view / projection / campos give the camera used
    to size the capacities. Gaussian tensors are read in place by every
    replay (update them in place to render changed Gaussians).
    """

    def __init__(self, settings: GaussianRasterizationSettings, means3D, opacities, scales,
                 rotations, dc=None, shs=None, colors_precomp=None,
                 options: RasterOptions | None = None, headroom: float = 1.25):
        dev = means3D.device
        self.device = dev
        self.opts = options or RasterOptions(schedule=_lib.SCHEDULE_FAST)
        self.copts = self.opts.to_c(False)
        self.copts.forward_only = 1  # rendering only: no bucket checkpoints
        self.P = P = int(means3D.shape[0])
        self.prefix_status = _prefix_lists(self.copts, self.opts, P, 1, dev)
        self.W, self.H = int(settings.image_width), int(settings.image_height)
        self.T = _lib.tiles_of(self.W, self.H, self.opts.tile_w)
        self.headroom = float(headroom)
        self.means3D = _in_place("means3D", means3D, dev, 3).reshape(P, 3)
        self.opacities = _in_place("opacities", opacities, dev, 1).reshape(P)
        self.scales = _in_place("scales", scales, dev, 3).reshape(P, 3)
        self.rotations = _in_place("rotations", rotations, dev, 4).reshape(P, 4)
        self.colors = self.dc = self.shs = None
        M = 0
        if colors_precomp is not None and colors_precomp.numel() > 0:
            self.colors = _in_place("colors_precomp", colors_precomp, dev, 3).reshape(P, 3)
        else:
            if dc is None:
                raise ValueError("dc is required when colours come from SH")
            self.dc = _in_place("dc", dc, dev, 3).reshape(P, 3)
            if shs is not None and shs.numel() > 0:
                self.shs = _in_place("shs", shs, dev, 3).reshape(P, -1, 3)
                M = int(self.shs.shape[1])
        # camera buffers the graph reads: view 16 | proj 16 | campos 3
        self.cambuf = torch.zeros(35, dtype=torch.float32, device=dev)
        bg = settings.bg if torch.is_tensor(settings.bg) else torch.tensor(settings.bg)
        self.bg = _f32(bg, dev).reshape(3)
        self.cam = Camera(self.W, self.H, float(settings.tanfovx), float(settings.tanfovy),
                          float(settings.scale_modifier), float(settings.mult),
                          int(settings.sh_degree), ctypes.c_void_p(self.cambuf.data_ptr()),
                          ctypes.c_void_p(self.cambuf.data_ptr() + 64),
                          ctypes.c_void_p(self.cambuf.data_ptr() + 128), _ptr(self.bg))
        self.gauss = Gaussians(P, M, _ptr(self.means3D), _ptr(self.opacities), _ptr(self.scales),
                               _ptr(self.rotations), _ptr(self.dc), _ptr(self.shs),
                               _ptr(self.colors), None)
        L = lib()
        self.gws_bytes = L.fgs_gauss_ws_size(P, self.W, self.H, ctypes.byref(self.copts))
        self.gws = torch.empty(self.gws_bytes, dtype=torch.uint8, device=dev)
        self.radii = torch.zeros(P, dtype=torch.int32, device=dev)
        self.out_color = torch.empty((3, self.H, self.W), dtype=torch.float32, device=dev)
        self.T_final = torch.empty((self.H, self.W, 1), dtype=torch.float32, device=dev)
        self.last_contributor = torch.empty((self.H, self.W, 1), dtype=torch.int32, device=dev)
        # device counts [R, NB capacity] inside the gauss workspace
        v = FrameView()
        check(L.fgs_frame_view_get(_ptr(self.gws), None, P, self.W, self.H, 0, 0,
                                   ctypes.byref(self.copts), ctypes.byref(v)), "fgs_frame_view_get")
        off = v.counts - self.gws.data_ptr()
        self.counts = self.gws[off:off + 8].view(torch.int32)
        self.max_counts = torch.zeros(2, dtype=torch.int32, device=dev)
        self.graph = None
        self.R_cap = self.NB_cap = 0
        self.set_camera(settings.viewmatrix, settings.projmatrix, settings.campos)
        R, NB = self._eager_counts()
        self._capture(R, NB)

    # -- camera -------------------------------------------------------------
    def set_camera(self, viewmatrix, projmatrix, campos):
        """Copy a camera into the buffers the graph reads (stream-ordered)."""
        self.cambuf[0:16].copy_(torch.as_tensor(viewmatrix).reshape(16))
        self.cambuf[16:32].copy_(torch.as_tensor(projmatrix).reshape(16))
        self.cambuf[32:35].copy_(torch.as_tensor(campos).reshape(3))

    # -- capture ------------------------------------------------------------
    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _preprocess(self):
        check(lib().fgs_preprocess(ctypes.byref(self.cam), ctypes.byref(self.gauss),
                                   ctypes.byref(self.copts), _ptr(self.gws), self.gws_bytes,
                                   _ptr(self.radii), self._stream()), "fgs_preprocess")

    def _eager_counts(self):
        self._preprocess()
        c = (ctypes.c_int64 * 2)()
        check(lib().fgs_read_counts(_ptr(self.gws), self.P, self.W, self.H,
                                    ctypes.byref(self.copts), c, self._stream()),
              "fgs_read_counts")
        return int(c[0]), int(c[1])

    def _frame(self):
        self._preprocess()
        check(lib().fgs_render(ctypes.byref(self.cam), ctypes.byref(self.gauss),
                               ctypes.byref(self.copts), _ptr(self.gws), self.gws_bytes,
                               _ptr(self.pws), self.pws.numel(), self.R_cap, self.NB_cap,
                               _ptr(self.out_color), _ptr(self.T_final),
                               _ptr(self.last_contributor), self._stream()), "fgs_render")
        torch.maximum(self.max_counts, self.counts[0:2], out=self.max_counts)

    def _capture(self, R: int, NB: int):
        B = self.opts.bucket_size
        self.R_cap = max(1024, int(math.ceil(R * self.headroom)))
        self.NB_cap = max(NB, (self.R_cap + B - 1) // B + self.T)
        n = lib().fgs_pair_ws_size(self.P, self.W, self.H, self.R_cap, self.NB_cap,
                                   ctypes.byref(self.copts))
        self.graph = None
        self.pws = torch.empty(n, dtype=torch.uint8, device=self.device)
        # one eager frame on a side stream first (allocator and module warm-up)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self._frame()
        torch.cuda.current_stream(self.device).wait_stream(side)
        self._prefix_check()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._frame()
        self.graph = g
        self.max_counts.zero_()

    # -- replay -------------------------------------------------------------
    def render(self, viewmatrix=None, projmatrix=None, campos=None) -> torch.Tensor:
        """Render one frame (optionally setting the camera first). Returns the
        (3, H, W) colour buffer, overwritten by the next replay."""
        if viewmatrix is not None:
            self.set_camera(viewmatrix, projmatrix, campos)
        self.graph.replay()
        return self.out_color

    def capacity(self) -> tuple[int, int]:
        return self.R_cap, self.NB_cap

    def prefix_redone(self) -> bool:
        """Whether some replay since the last call needed the whole lists
        (its prefixes did not suffice; its outputs are exact either way):
        then a capture with a smaller list_prefix_div saves that work."""
        return _redone(self.prefix_status)

    def check(self, grow: bool = False) -> bool:
        """True when every frame since the last check fitted the capacity.
        Otherwise raise CapacityError, or with grow=True re-capture with the
        largest counts seen (times the headroom) and return False; the frames
        that overflowed must then be rendered again."""
        R, NB = (int(x) for x in self.max_counts.tolist())
        self.max_counts.zero_()
        if R <= self.R_cap and NB <= self.NB_cap:
            return True
        if not grow:
            raise CapacityError(f"a frame needed {R} pairs / {NB} buckets, capacity is "
                                f"{self.R_cap} / {self.NB_cap}")
        self._capture(R, NB)
        return False


class CapturedTrainingStep(_PrefixCapture):
    """One training step of the rasterizer captured into one HIP graph.

    The step renders a batch of views (one view is a batch of one; uneven
    image sizes allowed, FAST schedule, one launch per stage) at fixed pair
    and bucket capacities, forms the gradient of the L2 loss
    sum_v mean((image_v - target_v)^2) against the views' target buffers,
    runs the rasterizer backward (raster adjoint, per-Gaussian adjoint summed
    over the views) and, when statistics buffers are given, the
    densification statistics. A replay reads in place whatever the caller
    last wrote into the Gaussian tensors (an optimizer updates them in
    place), the camera buffers (``set_cameras``) and the target buffers
    (``targets``), and overwrites ``grads`` and the output buffers; nothing
    in it waits for the host. Counts stay on the device: every replay folds
    them into a running maximum, and ``check()`` reports a replay that needed
    more pairs or buckets than the capacity (pairs beyond it are dropped,
    never written out of bounds; re-capture with ``grow=True``).
    """

    GRAD_NAMES = ("mean2D", "conic", "opacities", "colors", "means3D", "cov3D", "dc", "shs",
                  "scales", "rotations")

    def __init__(self, settings_list, means3D, opacities, scales, rotations, dc=None, shs=None,
                 colors_precomp=None, options: RasterOptions | None = None,
                 xyz_gradient_accum=None, xyz_gradient_accum_abs=None, max_radii=None,
                 denom=None, headroom: float = 1.25):
        from ._lib import Batch, Grads, MAX_VIEWS
        if not 1 <= len(settings_list) <= MAX_VIEWS:
            raise ValueError(f"a batch holds 1 to {MAX_VIEWS} views")
        dev = means3D.device
        self.device = dev
        self.opts = options or RasterOptions(schedule=_lib.SCHEDULE_FAST)
        if self.opts.schedule != _lib.SCHEDULE_FAST:
            raise ValueError("captured training steps use the FAST schedule")
        self.stats = None
        if xyz_gradient_accum is not None and xyz_gradient_accum.numel() > 0:
            self.stats = (xyz_gradient_accum, xyz_gradient_accum_abs, max_radii, denom)
        want_abs = xyz_gradient_accum_abs is not None and xyz_gradient_accum_abs.numel() > 0
        self.copts = self.opts.to_c(want_abs)
        self.P = P = int(means3D.shape[0])
        self.V = V = len(settings_list)
        self.prefix_status = _prefix_lists(self.copts, self.opts, P, V, dev)
        self.headroom = float(headroom)
        # the Gaussian tensors are read in place by every replay
        for name, t, per in (("means3D", means3D, 3), ("opacities", opacities, 1),
                             ("scales", scales, 3), ("rotations", rotations, 4)):
            _in_place(name, t, dev, per)
        self.means3D, self.opacities, self.scales, self.rotations = means3D, opacities, scales, rotations
        self.dc = self.shs = self.colors = None
        self.M = 0
        if colors_precomp is not None and colors_precomp.numel() > 0:
            self.colors = _in_place("colors_precomp", colors_precomp, dev, 3)
        else:
            if dc is None:
                raise ValueError("dc is required when colours come from SH")
            self.dc = _in_place("dc", dc, dev, 3)
            if shs is not None and shs.numel() > 0:
                self.shs = _in_place("shs", shs, dev, 3)
                self.M = int(shs.reshape(P, -1, 3).shape[1])
        self.sizes = [(int(s.image_height), int(s.image_width)) for s in settings_list]
        self.pix_off = [0]
        for H, W in self.sizes:
            self.pix_off.append(self.pix_off[-1] + H * W)
        HW = self.pix_off[-1]
        self.T = sum(_lib.tiles_of(W, H, self.opts.tile_w) for H, W in self.sizes)
        # per-view camera buffers the graph reads: view 16 | proj 16 | campos 3 | bg 3
        self.cambuf = torch.zeros(V, 40, dtype=torch.float32, device=dev)
        cams = (Camera * V)()
        base = self.cambuf.data_ptr()
        for v, st in enumerate(settings_list):
            row = base + v * 160
            bg = st.bg if torch.is_tensor(st.bg) else torch.tensor(st.bg)
            self.cambuf[v, 35:38].copy_(_f32(bg, dev).reshape(3))
            cams[v] = Camera(int(st.image_width), int(st.image_height), float(st.tanfovx),
                             float(st.tanfovy), float(st.scale_modifier), float(st.mult),
                             int(st.sh_degree), ctypes.c_void_p(row), ctypes.c_void_p(row + 64),
                             ctypes.c_void_p(row + 128), ctypes.c_void_p(row + 140))
        self.cams = cams
        self.batch = Batch(V, ctypes.cast(cams, ctypes.POINTER(Camera)))
        self.gauss = Gaussians(P, self.M, _ptr(means3D), _ptr(opacities), _ptr(scales),
                               _ptr(rotations), _ptr(self.dc), _ptr(self.shs), _ptr(self.colors),
                               None)
        self.set_cameras([(s.viewmatrix, s.projmatrix, s.campos) for s in settings_list])
        L = lib()
        self.gws_bytes = L.fgs_batch_gauss_ws_size(P, ctypes.byref(self.batch),
                                                   ctypes.byref(self.copts))
        self.gws = torch.empty(self.gws_bytes, dtype=torch.uint8, device=dev)
        self.radii = torch.zeros(V, P, dtype=torch.int32, device=dev)
        self.out_color = torch.empty(3 * HW, dtype=torch.float32, device=dev)
        self.T_final = torch.empty(HW, dtype=torch.float32, device=dev)
        self.last_contributor = torch.empty(HW, dtype=torch.int32, device=dev)
        # L2 loss: dL/dimage_v = 2 (image_v - target_v) / (3 H_v W_v)
        self.target = torch.zeros(3 * HW, dtype=torch.float32, device=dev)
        self.dimg = torch.empty(3 * HW, dtype=torch.float32, device=dev)
        self.wscale = torch.empty(3 * HW, dtype=torch.float32, device=dev)
        for o, (H, W) in zip(self.pix_off, self.sizes):
            self.wscale[3 * o:3 * o + 3 * H * W] = 2.0 / (3 * H * W)
        e = lambda *s: torch.empty(s, dtype=torch.float32, device=dev)
        self.grads = {"mean2D": e(V, P, 4), "conic": e(V, P, 2, 2), "opacities": e(P, 1),
                      "colors": e(P, 3), "means3D": e(P, 3), "cov3D": e(P, 6), "dc": e(P, 1, 3),
                      "shs": e(P, self.M, 3), "scales": e(P, 3), "rotations": e(P, 4)}
        self.cgrads = Grads(*[_ptr(self.grads[k]) for k in self.GRAD_NAMES])
        self.grad_ws = torch.empty(L.fgs_batch_grad_ws_size(P, V), dtype=torch.uint8, device=dev)
        # every replay's compositing zeroes the backward's raw sums; the
        # backward then skips its own zeroing
        self.copts.zero_grad_ws = self.grad_ws.data_ptr()
        self.copts.grad_ws_zeroed = 1
        v = FrameView()
        check(L.fgs_batch_frame_view_get(_ptr(self.gws), None, P, ctypes.byref(self.batch), 0, 0,
                                         ctypes.byref(self.copts), ctypes.byref(v)),
              "fgs_batch_frame_view_get")
        off = v.counts - self.gws.data_ptr()
        self.counts = self.gws[off:off + 8].view(torch.int32)
        self.max_counts = torch.zeros(2, dtype=torch.int32, device=dev)
        self.graph = None
        self.R_cap = self.NB_cap = 0
        R, NB = self._eager_counts()
        self._capture(R, NB)

    # -- inputs -------------------------------------------------------------
    def set_cameras(self, cameras):
        """Copy (viewmatrix, projmatrix, campos) per view into the buffers the
        graph reads (stream-ordered)."""
        if len(cameras) != self.V:
            raise ValueError(f"expected {self.V} cameras")
        for v, (view, proj, campos) in enumerate(cameras):
            self.cambuf[v, 0:16].copy_(torch.as_tensor(view).reshape(16))
            self.cambuf[v, 16:32].copy_(torch.as_tensor(proj).reshape(16))
            self.cambuf[v, 32:35].copy_(torch.as_tensor(campos).reshape(3))

    def targets(self) -> list:
        """Per-view (3, H, W) views of the target buffer (write targets here)."""
        return [self.target[3 * o:3 * o + 3 * H * W].view(3, H, W)
                for o, (H, W) in zip(self.pix_off, self.sizes)]

    def images(self) -> list:
        """Per-view (3, H, W) views of the colour buffer of the last replay."""
        return [self.out_color[3 * o:3 * o + 3 * H * W].view(3, H, W)
                for o, (H, W) in zip(self.pix_off, self.sizes)]

    # -- capture ------------------------------------------------------------
    def _s(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _preprocess(self):
        check(lib().fgs_batch_preprocess(ctypes.byref(self.batch), ctypes.byref(self.gauss),
                                         ctypes.byref(self.copts), _ptr(self.gws), self.gws_bytes,
                                         _ptr(self.radii), self._s()), "fgs_batch_preprocess")

    def _eager_counts(self):
        self._preprocess()
        c = (ctypes.c_int64 * 2)()
        check(lib().fgs_batch_read_counts(_ptr(self.gws), self.P, ctypes.byref(self.batch),
                                          ctypes.byref(self.copts), c, self._s()),
              "fgs_batch_read_counts")
        return int(c[0]), int(c[1])

    def _step(self):
        L = lib()
        self._preprocess()
        check(L.fgs_batch_render(ctypes.byref(self.batch), ctypes.byref(self.gauss),
                                 ctypes.byref(self.copts), _ptr(self.gws), self.gws_bytes,
                                 _ptr(self.pws), self.pws.numel(), self.R_cap, self.NB_cap,
                                 _ptr(self.out_color), _ptr(self.T_final),
                                 _ptr(self.last_contributor), self._s()), "fgs_batch_render")
        torch.maximum(self.max_counts, self.counts[0:2], out=self.max_counts)
        torch.sub(self.out_color, self.target, out=self.dimg)
        self.dimg.mul_(self.wscale)
        check(L.fgs_batch_backward(ctypes.byref(self.batch), ctypes.byref(self.gauss),
                                   ctypes.byref(self.copts), _ptr(self.gws), _ptr(self.pws),
                                   self.R_cap, self.NB_cap, _ptr(self.radii), _ptr(self.T_final),
                                   _ptr(self.last_contributor), _ptr(self.dimg),
                                   _ptr(self.grad_ws), ctypes.byref(self.cgrads), self._s()),
              "fgs_batch_backward")
        if self.stats is not None:
            accum, accum_abs, max_radii, denom = self.stats
            check(L.fgs_batch_densify_stats(self.P, self.V, _ptr(self.radii),
                                            _ptr(self.grads["mean2D"]), _ptr(accum),
                                            _ptr(accum_abs), _ptr(max_radii), _ptr(denom),
                                            self._s()), "fgs_batch_densify_stats")

    def _capture(self, R: int, NB: int):
        B = self.opts.bucket_size
        self.R_cap = max(1024, int(math.ceil(R * self.headroom)))
        self.NB_cap = max(NB, (self.R_cap + B - 1) // B + self.T)
        n = lib().fgs_batch_pair_ws_size(self.P, ctypes.byref(self.batch), self.R_cap,
                                         self.NB_cap, ctypes.byref(self.copts))
        self.graph = None
        self.pws = torch.empty(n, dtype=torch.uint8, device=self.device)
        # capturing updates no statistics: they are saved and restored around
        # the warm-up step
        saved = None if self.stats is None else [None if t is None else t.clone()
                                                 for t in self.stats]
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self._step()
        torch.cuda.current_stream(self.device).wait_stream(side)
        self._prefix_check()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._step()
        self.graph = g
        if saved is not None:
            for t, s in zip(self.stats, saved):
                if t is not None:
                    t.copy_(s)
        self.max_counts.zero_()

    # -- replay -------------------------------------------------------------
    def step(self) -> dict:
        """Replay the step; returns the gradient buffers (overwritten by the
        next replay): parameter gradients summed over the views, screen-space
        ones (mean2D, conic) per view."""
        self.graph.replay()
        return self.grads

    def capacity(self) -> tuple[int, int]:
        return self.R_cap, self.NB_cap

    def prefix_redone(self) -> bool:
        """Whether some replay since the last call needed the whole lists
        (its prefixes did not suffice; its outputs are exact either way):
        then a capture with a smaller list_prefix_div saves that work."""
        return _redone(self.prefix_status)

    def check(self, grow: bool = False) -> bool:
        """True when every replay since the last check fitted the capacity.
        Otherwise raise CapacityError, or with grow=True re-capture with the
        largest counts seen (times the headroom) and return False; the steps
        that overflowed must then be run again."""
        R, NB = (int(x) for x in self.max_counts.tolist())
        self.max_counts.zero_()
        if R <= self.R_cap and NB <= self.NB_cap:
            return True
        if not grow:
            raise CapacityError(f"a step needed {R} pairs / {NB} buckets, capacity is "
                                f"{self.R_cap} / {self.NB_cap}")
        self._capture(R, NB)
        return False
