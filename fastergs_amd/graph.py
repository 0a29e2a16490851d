"""Forward rendering captured into one HIP graph per frame.

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


class CapturedRenderer:
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
        self.P = P = int(means3D.shape[0])
        self.W, self.H = int(settings.image_width), int(settings.image_height)
        self.T = ((self.W + 15) // 16) * ((self.H + 15) // 16)
        self.headroom = float(headroom)
        self.means3D = _f32(means3D, dev).reshape(P, 3)
        self.opacities = _f32(opacities, dev).reshape(P)
        self.scales = _f32(scales, dev).reshape(P, 3)
        self.rotations = _f32(rotations, dev).reshape(P, 4)
        self.colors = self.dc = self.shs = None
        M = 0
        if colors_precomp is not None and colors_precomp.numel() > 0:
            self.colors = _f32(colors_precomp, dev).reshape(P, 3)
        else:
            self.dc = _f32(dc, dev).reshape(P, 3)
            if shs is not None and shs.numel() > 0:
                self.shs = _f32(shs, dev).reshape(P, -1, 3)
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
