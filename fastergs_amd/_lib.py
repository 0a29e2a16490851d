"""ctypes binding of libfastergs.so (declarations in include/fastergs.h).

The library must be built in-tree (``python fastergs_amd/build.py``);
importing the rasterizer without it raises immediately -- there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfastergs.so")

SCHEDULE_PLAIN = 0
SCHEDULE_FAST = 1

TILES_RASTER = 0
TILES_HEURISTIC = 1
TILES_LEARNED = 2
MLP_IN, MLP_HIDDEN, MLP_OUT = 8, 64, 2
MLP_NPARAMS = MLP_HIDDEN * MLP_IN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN \
    + MLP_OUT * MLP_HIDDEN + MLP_OUT  # 4866

vp = ctypes.c_void_p


class Camera(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int),
                ("tanfovx", ctypes.c_float), ("tanfovy", ctypes.c_float),
                ("scale_modifier", ctypes.c_float), ("mult", ctypes.c_float),
                ("sh_degree", ctypes.c_int),
                ("viewmatrix", vp), ("projmatrix", vp), ("campos", vp), ("bg", vp)]


class Gaussians(ctypes.Structure):
    _fields_ = [("P", ctypes.c_int), ("max_coeffs", ctypes.c_int),
                ("means3D", vp), ("opacities", vp), ("scales", vp), ("rotations", vp),
                ("dc", vp), ("shs", vp), ("colors_precomp", vp), ("culling", vp)]


class TileSched(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("split_len", ctypes.c_int), ("mlp_params", vp),
                ("features", ctypes.c_int), ("log_cycles", ctypes.c_int), ("order", ctypes.c_int),
                ("items", vp), ("n_items", vp), ("prev_features", vp), ("features_out", vp)]


def tiles_of(width: int, height: int, tile_w: int = 16) -> int:
    """Tiles of a width x height image on tile_w x (256 / tile_w) tiles."""
    tw = tile_w or 16
    th = 256 // tw
    return ((width + tw - 1) // tw) * ((height + th - 1) // th)


class Options(ctypes.Structure):
    _fields_ = [("schedule", ctypes.c_int), ("bucket_size", ctypes.c_int),
                ("split_tiles", ctypes.c_int), ("abs_grad", ctypes.c_int),
                ("tile_sched", ctypes.POINTER(TileSched)), ("forward_only", ctypes.c_int),
                ("composite_queue", ctypes.c_int), ("tile_w", ctypes.c_int),
                ("list_prefix", ctypes.c_int), ("zero_grad_ws", vp), ("grad_ws_zeroed", ctypes.c_int),
                ("list_status", vp)]


class AdamTensor(ctypes.Structure):
    _fields_ = [("param", vp), ("grad", vp), ("exp_avg", vp), ("exp_avg_sq", vp),
                ("M", ctypes.c_int), ("lr", ctypes.c_float), ("eps", ctypes.c_float)]


class Batch(ctypes.Structure):
    _fields_ = [("n_views", ctypes.c_int), ("cameras", ctypes.POINTER(Camera))]


MAX_VIEWS = 32


class FrameView(ctypes.Structure):
    _fields_ = [(n, vp) for n in (
        "geom", "splat", "pair_count", "tile_count", "ranges", "bucket_off", "max_contrib",
        "counts", "sorted_keys", "sorted_vals", "bucket_ranges", "bucket_tile", "acc_t", "acc_c",
        "pixel_colors", "tile_features", "tile_pred", "work_items", "item_ticks")]


class Grads(ctypes.Structure):
    _fields_ = [(n, vp) for n in (
        "dmean2D", "dconic", "dopacity", "dcolor", "dmean3D", "dcov3D", "ddc", "dsh", "dscale",
        "drot", "xyz_grad_accum", "xyz_grad_accum_abs", "max_radii", "denom")]


# name -> (restype, argtypes); the list is the ABI surface checked by tests.
SIGNATURES = {
    "fgs_gauss_ws_size": (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(Options)]),
    "fgs_pair_ws_size": (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int64, ctypes.c_int64,
                                           ctypes.POINTER(Options)]),
    "fgs_grad_ws_size": (ctypes.c_size_t, [ctypes.c_int]),
    "fgs_preprocess": (ctypes.c_int, [ctypes.POINTER(Camera), ctypes.POINTER(Gaussians),
                                      ctypes.POINTER(Options), vp, ctypes.c_size_t, vp, vp]),
    "fgs_read_counts": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(Options), ctypes.POINTER(ctypes.c_int64),
                                       vp]),
    "fgs_render": (ctypes.c_int, [ctypes.POINTER(Camera), ctypes.POINTER(Gaussians),
                                  ctypes.POINTER(Options), vp, ctypes.c_size_t, vp,
                                  ctypes.c_size_t, ctypes.c_int64, ctypes.c_int64, vp, vp, vp,
                                  vp]),
    "fgs_backward": (ctypes.c_int, [ctypes.POINTER(Camera), ctypes.POINTER(Gaussians),
                                    ctypes.POINTER(Options), vp, vp, ctypes.c_int64,
                                    ctypes.c_int64, vp, vp, vp, vp, vp, ctypes.POINTER(Grads),
                                    vp]),
    "fgs_densify_stats": (ctypes.c_int, [ctypes.c_int, vp, vp, vp, vp, vp, vp, vp]),
    "fgs_rcp_selftest": (ctypes.c_int, [vp, vp]),
    "fgs_fma_probe": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, vp]),
    "fgs_frame_view_get": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int64, ctypes.c_int64,
                                          ctypes.POINTER(Options), ctypes.POINTER(FrameView)]),
    "fgs_bucket_tables": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(Options),
                                         vp]),
    "fgs_sort_ws_size": (ctypes.c_size_t, [ctypes.c_int64]),
    "fgs_sort_pairs": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int,
                                      ctypes.c_int, vp, ctypes.c_size_t, vp]),
    "fgs_sort_check": (ctypes.c_int, [vp, ctypes.c_int64, vp]),
    "fgs_sort_pairs_noalloc": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int,
                                              ctypes.c_int, vp]),
    "fgs_scan_ws_size": (ctypes.c_size_t, [ctypes.c_int64]),
    "fgs_scan_u32": (ctypes.c_int, [vp, vp, ctypes.c_int64, vp, vp, ctypes.c_size_t, vp]),
    "fgs_mlp_forward": (ctypes.c_int, [vp, vp, ctypes.c_int64, vp, ctypes.c_int, vp]),
    "fgs_mlp_grad_ws_size": (ctypes.c_size_t, [ctypes.c_int64]),
    "fgs_mlp_loss_grad": (ctypes.c_int, [vp, vp, vp, ctypes.c_int64, ctypes.c_float, vp, vp, vp,
                                         ctypes.c_size_t, vp]),
    "fgs_adam_update": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_int64, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_int, vp]),
    "fgs_batch_gauss_ws_size": (ctypes.c_size_t, [ctypes.c_int, ctypes.POINTER(Batch),
                                                  ctypes.POINTER(Options)]),
    "fgs_batch_pair_ws_size": (ctypes.c_size_t, [ctypes.c_int, ctypes.POINTER(Batch),
                                                 ctypes.c_int64, ctypes.c_int64,
                                                 ctypes.POINTER(Options)]),
    "fgs_batch_grad_ws_size": (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int]),
    "fgs_batch_preprocess": (ctypes.c_int, [ctypes.POINTER(Batch), ctypes.POINTER(Gaussians),
                                            ctypes.POINTER(Options), vp, ctypes.c_size_t, vp,
                                            vp]),
    "fgs_batch_read_counts": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(Batch),
                                             ctypes.POINTER(Options),
                                             ctypes.POINTER(ctypes.c_int64), vp]),
    "fgs_batch_render": (ctypes.c_int, [ctypes.POINTER(Batch), ctypes.POINTER(Gaussians),
                                        ctypes.POINTER(Options), vp, ctypes.c_size_t, vp,
                                        ctypes.c_size_t, ctypes.c_int64, ctypes.c_int64, vp, vp,
                                        vp, vp]),
    "fgs_batch_backward": (ctypes.c_int, [ctypes.POINTER(Batch), ctypes.POINTER(Gaussians),
                                          ctypes.POINTER(Options), vp, vp, ctypes.c_int64,
                                          ctypes.c_int64, vp, vp, vp, vp, vp,
                                          ctypes.POINTER(Grads), vp]),
    "fgs_batch_densify_stats": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp,
                                               vp, vp]),
    "fgs_batch_frame_view_get": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.POINTER(Batch),
                                                ctypes.c_int64, ctypes.c_int64,
                                                ctypes.POINTER(Options),
                                                ctypes.POINTER(FrameView)]),
    "fgs_batch_bucket_tables": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.POINTER(Batch),
                                               ctypes.c_int64, ctypes.c_int64,
                                               ctypes.POINTER(Options), vp]),
    "fgs_sparse_adam": (ctypes.c_int, [ctypes.POINTER(AdamTensor), ctypes.c_int, vp, ctypes.c_int,
                                       ctypes.c_float, ctypes.c_float, vp]),
    "fgs_ssim_forward": (ctypes.c_int, [ctypes.c_float, ctypes.c_float, vp, vp, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp,
                                        vp]),
    "fgs_ssim_backward": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, vp, vp]),
    "fgs_render_simp": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(Options),
                                       vp, vp, vp, vp]),
    "fgs_render_depth": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(Options)]
                         + [vp] * 11),
    "fgs_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "fgs_profile_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "fgs_last_error": (ctypes.c_char_p, []),
    "fgs_build_info": (ctypes.c_char_p, []),
}

_lib = None


class FasterGSError(RuntimeError):
    pass


def lib():
    """Load libfastergs.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python fastergs_amd/build.py` "
                "(the rasterizer has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().fgs_last_error().decode(errors="replace")
        raise FasterGSError(f"{what} failed ({rc}): {msg}")


def profile_enable(on: bool = True):
    lib().fgs_profile_enable(1 if on else 0)


def profile_read() -> dict:
    """{kernel name: (total ms, launches)} since the last profile_enable()."""
    L = lib()
    n = L.fgs_profile_read(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    L.fgs_profile_read(buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        name, ms, cnt = line.split()
        out[name] = (float(ms), int(cnt))
    return out
