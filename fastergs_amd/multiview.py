"""Multi-view training steps: several cameras, one launch per stage.

The FasterGS training loops render one camera per step
(3DGS/train_faster.py); a batch of views of uneven resolutions is rendered
here by the FAST schedule in one pass per stage (preprocess, emission, sort,
ranges, buckets, compositing, backward, preprocess adjoint) through the
library's ``fgs_batch_*`` entry points:

    images, radii = rasterize_views(settings_list, means3D, opacities, dc=dc,
                                    shs=shs, scales=scales, rotations=rotations)
    loss = sum(((img - tgt) ** 2).mean() for img, tgt in zip(images, targets))
    loss.backward()   # gradients summed over the views

Each view's image, radii and screen-space gradients equal those of a
single-view call with its settings; the parameter gradients are the sum over
the views (added in view order), and the densification statistics, when
given, receive every view's contribution in view order (like consecutive
single-view backward passes).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import Batch, Camera, FrameView, Gaussians, Grads, check, lib
from .rasterizer import RasterOptions, _Grads, _PrefixPolicy, _f32, _prefix_policies, _ptr, _stream


class BatchFrame:
    """Forward evaluation of one batch of views: inputs, workspaces, outputs."""

    def __init__(self, settings_list, means3D, opacities, scales, rotations, dc=None, shs=None,
                 colors_precomp=None, options: RasterOptions | None = None,
                 abs_grad: bool = True, forward_only: bool = False):
        if not 1 <= len(settings_list) <= _lib.MAX_VIEWS:
            raise ValueError(f"a batch holds 1 to {_lib.MAX_VIEWS} views")
        dev = means3D.device
        self.device = dev
        self.opts = options or RasterOptions(schedule=_lib.SCHEDULE_FAST)
        if self.opts.schedule != _lib.SCHEDULE_FAST:
            raise ValueError("batches of views use the FAST schedule")
        self.copts = self.opts.to_c(abs_grad)
        self.forward_only = bool(forward_only)  # no bucket checkpoints: not differentiable
        self.copts.forward_only = 1 if forward_only else 0
        self.P = P = int(means3D.shape[0])
        self.V = V = len(settings_list)
        self.means3D = _f32(means3D, dev).reshape(P, 3)
        self.opacities = _f32(opacities, dev).reshape(P)
        self.scales = _f32(scales, dev).reshape(P, 3)
        self.rotations = _f32(rotations, dev).reshape(P, 4)
        self.colors = self.dc = self.shs = None
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
        self.sizes = [(int(s.image_height), int(s.image_width)) for s in settings_list]
        self.pix_off = [0]
        for H, W in self.sizes:
            self.pix_off.append(self.pix_off[-1] + H * W)
        keep = []
        cams = (Camera * V)()
        for v, st in enumerate(settings_list):
            view = _f32(st.viewmatrix, dev).reshape(16)
            proj = _f32(st.projmatrix, dev).reshape(16)
            campos = _f32(st.campos, dev).reshape(3)
            bg = st.bg if torch.is_tensor(st.bg) else torch.tensor(st.bg)
            bg = _f32(bg, dev).reshape(3)
            keep += [view, proj, campos, bg]
            cams[v] = Camera(int(st.image_width), int(st.image_height), float(st.tanfovx),
                             float(st.tanfovy), float(st.scale_modifier), float(st.mult),
                             int(st.sh_degree), _ptr(view), _ptr(proj), _ptr(campos), _ptr(bg))
        self._keep = keep
        self.cams = cams
        self.batch = Batch(V, ctypes.cast(cams, ctypes.POINTER(Camera)))
        self.gauss = Gaussians(P, self.M, _ptr(self.means3D), _ptr(self.opacities),
                               _ptr(self.scales), _ptr(self.rotations), _ptr(self.dc),
                               _ptr(self.shs), _ptr(self.colors), None)
        self.R = self.NB = 0

    def _s(self):
        return ctypes.c_void_p(_stream(self.device))

    def preprocess(self):
        L = lib()
        self.radii = torch.empty(self.V, self.P, dtype=torch.int32, device=self.device)  # every entry written
        n = L.fgs_batch_gauss_ws_size(self.P, ctypes.byref(self.batch), ctypes.byref(self.copts))
        self.gws = torch.empty(n, dtype=torch.uint8, device=self.device)
        check(L.fgs_batch_preprocess(ctypes.byref(self.batch), ctypes.byref(self.gauss),
                                     ctypes.byref(self.copts), _ptr(self.gws), n, _ptr(self.radii),
                                     self._s()), "fgs_batch_preprocess")
        c = (ctypes.c_int64 * 2)()
        check(L.fgs_batch_read_counts(_ptr(self.gws), self.P, ctypes.byref(self.batch),
                                      ctypes.byref(self.copts), c, self._s()),
              "fgs_batch_read_counts")
        self.R, self.NB = int(c[0]), int(c[1])
        return self

    def render(self):
        L, dev = lib(), self.device
        # prefix lists of the batch's nearest records first, sized per batch
        # shape by the drop-in API's policy (RasterOptions.list_prefix_div)
        pol, div = None, int(self.opts.list_prefix_div)
        NR = self.P * self.V
        if div > 1 and self.opts.tile_sched is None:
            key = (tuple(self.sizes), self.opts.bucket_size, self.opts.tile_w, dev, div)
            pol = _prefix_policies.setdefault(key, _PrefixPolicy(div))
            div = pol.divisor()
            if div > 1 and NR > div:
                status = pol.status_word()
                self.copts.list_prefix = -(-NR // div)
                self.copts.list_status = status.data_ptr()
            else:
                pol = None
        n = L.fgs_batch_pair_ws_size(self.P, ctypes.byref(self.batch), self.R, self.NB,
                                     ctypes.byref(self.copts))
        self.pws = torch.empty(n, dtype=torch.uint8, device=dev)
        HW = self.pix_off[-1]
        self.out_color = torch.empty(3 * HW, dtype=torch.float32, device=dev)
        self.T_final = torch.empty(HW, dtype=torch.float32, device=dev)
        self.last_contributor = torch.empty(HW, dtype=torch.int32, device=dev)
        check(L.fgs_batch_render(ctypes.byref(self.batch), ctypes.byref(self.gauss),
                                 ctypes.byref(self.copts), _ptr(self.gws), self.gws.numel(),
                                 _ptr(self.pws), n, self.R, self.NB, _ptr(self.out_color),
                                 _ptr(self.T_final), _ptr(self.last_contributor), self._s()),
              "fgs_batch_render")
        if pol is not None:
            pol.track(status, div)
            self.copts.list_status = None
            self.prefix_lists = True
        return self

    def complete_lists(self):
        """Whole per-tile lists for a batch rendered from prefix lists (the
        inspection views read whole lists); the outputs keep their values."""
        if not getattr(self, "prefix_lists", False):
            return self
        self.copts.list_prefix = 0
        check(lib().fgs_batch_render(ctypes.byref(self.batch), ctypes.byref(self.gauss),
                                     ctypes.byref(self.copts), _ptr(self.gws), self.gws.numel(),
                                     _ptr(self.pws), self.pws.numel(), self.R, self.NB,
                                     _ptr(self.out_color), _ptr(self.T_final),
                                     _ptr(self.last_contributor), self._s()), "fgs_batch_render")
        self.prefix_lists = False
        return self

    def forward(self):
        return self.preprocess().render()

    def images(self, flat: torch.Tensor | None = None) -> list:
        """Per-view (3, H, W) views of the concatenated colour buffer."""
        flat = self.out_color if flat is None else flat
        return [flat[3 * o:3 * o + 3 * H * W].view(3, H, W)
                for o, (H, W) in zip(self.pix_off, self.sizes)]

    def per_view(self, flat: torch.Tensor) -> list:
        """Per-view (H, W) views of a concatenated per-pixel buffer."""
        return [flat[o:o + H * W].view(H, W) for o, (H, W) in zip(self.pix_off, self.sizes)]

    def backward(self, grad_flat: torch.Tensor, skip=()) -> dict:
        """Gradients of the batch; `skip` may name "conic", "colors" and
        "cov3D", which are then neither computed into memory nor returned."""
        L, dev, P, V = lib(), self.device, self.P, self.V
        if self.forward_only:
            raise ValueError("the batch was rendered forward-only")
        if not set(skip) <= {"conic", "colors", "cov3D"}:
            raise ValueError("only conic, colors and cov3D gradients can be skipped")
        g = _f32(grad_flat, dev).reshape(-1)
        e = lambda *s: torch.empty(s, dtype=torch.float32, device=dev)
        shapes = {"mean2D": (V, P, 4), "conic": (V, P, 2, 2), "opacities": (P, 1),
                  "colors": (P, 3), "means3D": (P, 3), "cov3D": (P, 6), "dc": (P, 1, 3),
                  "shs": (P, self.M, 3), "scales": (P, 3), "rotations": (P, 4)}
        out = {k: e(*sh) for k, sh in shapes.items() if k not in skip}
        grads = Grads(*[_ptr(out.get(k)) for k in ("mean2D", "conic", "opacities", "colors",
                                                    "means3D", "cov3D", "dc", "shs", "scales",
                                                    "rotations")])
        gws = torch.empty(L.fgs_batch_grad_ws_size(P, V), dtype=torch.uint8, device=dev)
        check(L.fgs_batch_backward(ctypes.byref(self.batch), ctypes.byref(self.gauss),
                                   ctypes.byref(self.copts), _ptr(self.gws), _ptr(self.pws),
                                   self.R, self.NB, _ptr(self.radii), _ptr(self.T_final),
                                   _ptr(self.last_contributor), _ptr(g), _ptr(gws),
                                   ctypes.byref(grads), self._s()), "fgs_batch_backward")
        return _Grads(out, gws, V * P, (V, P, 12))

    def densify_stats(self, dmean2D, accum, accum_abs, max_radii, denom):
        check(lib().fgs_batch_densify_stats(self.P, self.V, _ptr(self.radii), _ptr(dmean2D),
                                            _ptr(accum), _ptr(accum_abs), _ptr(max_radii),
                                            _ptr(denom), self._s()), "fgs_batch_densify_stats")

    def view_arrays(self) -> dict:
        """Copies of the batch's intermediate arrays (lists, ranges, buckets)."""
        self.complete_lists()
        if getattr(self, "pws", None) is not None:
            # the bucket tables are built on demand (no pass of the step reads them)
            check(lib().fgs_batch_bucket_tables(_ptr(self.gws), _ptr(self.pws), self.P,
                                                ctypes.byref(self.batch), self.R, self.NB,
                                                ctypes.byref(self.copts), self._s()),
                  "fgs_batch_bucket_tables")
        v = FrameView()
        check(lib().fgs_batch_frame_view_get(_ptr(self.gws), _ptr(getattr(self, "pws", None)),
                                             self.P, ctypes.byref(self.batch), self.R, self.NB,
                                             ctypes.byref(self.copts), ctypes.byref(v)),
              "fgs_batch_frame_view_get")
        T = sum(_lib.tiles_of(W, H, self.opts.tile_w) for H, W in self.sizes)
        PV = self.P * self.V
        spec = {"splat": (torch.float32, PV * 12), "ranges": (torch.int32, 2 * T),
                "max_contrib": (torch.int32, T), "counts": (torch.int32, 16),
                "sorted_vals": (torch.int32, self.R)}
        res = {}
        item = {torch.float32: 4, torch.int32: 4}
        for name, (dt, n) in spec.items():
            addr = getattr(v, name)
            if not addr:
                continue
            for buf in (self.gws, getattr(self, "pws", None)):
                if buf is not None and buf.data_ptr() <= addr < buf.data_ptr() + buf.numel():
                    off = addr - buf.data_ptr()
                    res[name] = buf[off:off + n * item[dt]].view(dt).clone()
        return res


class _RasterizeViews(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means3D, dc, shs, colors_precomp, opacities, scales, rotations, frame_args,
                stats):
        settings_list, options, grad_enabled = frame_args
        want_abs = stats[1] is not None and stats[1].numel() > 0
        frame = BatchFrame(settings_list, means3D, opacities, scales, rotations, dc=dc, shs=shs,
                           colors_precomp=colors_precomp, options=options, abs_grad=want_abs,
                           forward_only=not grad_enabled or
                           not any(ctx.needs_input_grad[:7])).forward()
        ctx.frame = frame
        ctx.shapes = [t.shape if torch.is_tensor(t) else None for t in
                      (means3D, dc, shs, colors_precomp, opacities, scales, rotations)]
        ctx.stats = stats
        ctx.mark_non_differentiable(frame.radii)
        ctx.set_materialize_grads(False)  # no zero-filled gradient for the radii
        return frame.out_color, frame.radii

    @staticmethod
    def backward(ctx, grad_flat, _grad_radii):
        frame: BatchFrame = ctx.frame
        if grad_flat is None:  # the images took no part in the loss
            ctx.frame = None
            return (None,) * 9
        s = ctx.shapes
        # conic and cov3D gradients are never returned, colours only when precomputed
        skip = ["conic", "cov3D"]
        if s[3] is None or s[3].numel() == 0:
            skip.append("colors")
        g = frame.backward(grad_flat, skip=skip)
        accum, accum_abs, max_radii, denom = ctx.stats
        if accum is not None and accum.numel() != 0:
            frame.densify_stats(g["mean2D"], accum, accum_abs, max_radii, denom)
        ctx.frame = None

        def shaped(t, shape):
            if shape is None or t is None:  # input None, or not computed
                return None
            if t.numel() == shape.numel() and t.numel() > 0:
                return t.reshape(shape)
            return t

        return (shaped(g["means3D"], s[0]), shaped(g["dc"], s[1]), shaped(g["shs"], s[2]),
                shaped(g.get("colors"), s[3]), shaped(g["opacities"], s[4]),
                shaped(g["scales"], s[5]), shaped(g["rotations"], s[6]), None, None)


def rasterize_views(settings_list, means3D, opacities, dc=None, shs=None, colors_precomp=None,
                    scales=None, rotations=None, xyz_gradient_accum=None,
                    xyz_gradient_accum_abs=None, max_radii=None, denom=None,
                    options: RasterOptions | None = None):
    """Render a batch of views; returns (list of (3, H, W) images, radii (V, P))."""
    if (shs is None and colors_precomp is None) or (shs is not None and colors_precomp is not None):
        raise ValueError("Please provide excatly one of either SHs or precomputed colors!")
    if scales is None or rotations is None:
        raise ValueError("scales and rotations are required")
    opts = options or RasterOptions(schedule=_lib.SCHEDULE_FAST)
    stats = (xyz_gradient_accum, xyz_gradient_accum_abs, max_radii, denom)
    flat, radii = _RasterizeViews.apply(means3D, dc, shs, colors_precomp, opacities, scales,
                                        rotations,
                                        (list(settings_list), opts, torch.is_grad_enabled()),
                                        stats)
    out, o = [], 0
    for st in settings_list:
        H, W = int(st.image_height), int(st.image_width)
        out.append(flat[3 * o:3 * (o + H * W)].view(3, H, W))
        o += H * W
    return out, radii
