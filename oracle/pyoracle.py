"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg, always as the checker or the timed CPU
baseline, never as a product code path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i32, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        for suf in ("f32", "f64"):
            create = getattr(L, "orc_create_" + suf)
            create.restype = vp
            create.argtypes = [i32] + [vp] * 6 + [i32] + [vp] * 6 + [f64] * 4 + [i32] * 6
            getattr(L, "orc_forward_" + suf).argtypes = [vp]
            getattr(L, "orc_backward_" + suf).argtypes = [vp, vp]
            g = getattr(L, "orc_get_" + suf)
            g.argtypes = [vp, ctypes.c_char_p, vp]
            g.restype = ctypes.c_longlong
            getattr(L, "orc_destroy_" + suf).argtypes = [vp]
        L.orc_stats_f32.argtypes = [vp] * 5
        L.orc_gauss_grads_from_raw_f32.argtypes = [vp, vp]
        L.orc_exp2f.argtypes = [ctypes.c_float]
        L.orc_exp2f.restype = ctypes.c_float
        L.orc_log2f.argtypes = [ctypes.c_float]
        L.orc_log2f.restype = ctypes.c_float
        L.orc_tile_hit_f32.argtypes = [i32, i32] + [ctypes.c_float] * 6 + [i32]
        L.orc_tile_hit_f32.restype = i32
        L.orc_num_threads.restype = i32
        L.orc_tile_terms_check.argtypes = [ctypes.c_float] * 6 + [i32, i32, vp, i32]
        L.orc_tile_terms_check.restype = i32
        L.orc_render_simp_f32.argtypes = [vp, vp, vp, vp]
        L.orc_render_depth_f32.argtypes = [vp] * 6
        L.orc_inverse4_f64.argtypes = [vp, vp]
        L.orc_depth_ray_f64.argtypes = [vp, vp, i32, i32, i32, i32, vp]
        L.orc_depth_ray_f32.argtypes = [vp, vp, i32, i32, i32, i32, vp]
        L.orc_ellipsoid_f64.argtypes = [vp] * 6
        _lib = L
    return _lib


_DTYPES = {
    "radii": np.int32, "pair_count": np.uint32, "keys": np.uint64, "vals": np.uint32,
    "ranges": np.uint32, "bucket_off": np.uint32, "bucket_ranges": np.uint32,
    "bucket_tile": np.uint32, "contrib": np.int32, "max_contrib": np.uint32, "raw": np.float64,
}


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One forward(+backward) evaluation of the CPU oracle.

    mode 0 = bucketed compositing (reference semantics), 1 = sequential
    (plain path). dtype float32 (bit-exact checker) or float64 (FD checks).
    tile_w: tiles of tile_w x (256 / tile_w) pixels (16: the reference's).
    """

    def __init__(self, scene: dict, cam, bg=(0.0, 0.0, 0.0), sh_degree: int = 3,
                 bucket: int = 32, mode: int = 0, dtype=np.float32, scale_modifier: float = 1.0,
                 mult: float = 1.0, colors=None, culling=None, use_sh: bool = True,
                 tile_w: int = 16):
        self.L = lib()
        self.dtype = np.dtype(dtype)
        self.suf = "f32" if self.dtype == np.float32 else "f64"
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=self.dtype)
        self._keep = []
        means = c(scene["means3D"]).reshape(-1, 3)
        self.P = P = means.shape[0]
        opac = c(scene["opacities"]).reshape(-1)
        scales = c(scene["scales"]).reshape(-1, 3)
        rots = c(scene["rotations"]).reshape(-1, 4)
        dc = shs = None
        M = 0
        if colors is None and use_sh:
            dc = c(scene["dc"]).reshape(P, 3)
            shs = c(scene["shs"])
            M = shs.shape[1] if shs.ndim == 3 else shs.size // (3 * P)
        col = c(colors)
        cull = None if culling is None else np.ascontiguousarray(culling, dtype=np.uint8)
        view = c(cam.viewmatrix).reshape(16)
        proj = c(cam.projmatrix).reshape(16)
        campos = c(cam.campos).reshape(3)
        bgv = c(np.asarray(bg)).reshape(3)
        self._keep = [means, opac, scales, rots, dc, shs, col, cull, view, proj, campos, bgv]
        self.W, self.H = cam.width, cam.height
        self.tile_w, self.tile_h = tile_w, 256 // tile_w
        self.gx = (self.W + self.tile_w - 1) // self.tile_w
        self.gy = (self.H + self.tile_h - 1) // self.tile_h
        self.bucket = bucket
        self.h = getattr(self.L, "orc_create_" + self.suf)(
            P, _ptr(means), _ptr(opac), _ptr(scales), _ptr(rots), _ptr(dc), _ptr(shs), M,
            _ptr(col), _ptr(cull), _ptr(view), _ptr(proj), _ptr(campos), _ptr(bgv),
            float(cam.tanfovx), float(cam.tanfovy), float(scale_modifier), float(mult),
            self.W, self.H, sh_degree, bucket, mode, tile_w)

    def forward(self):
        getattr(self.L, "orc_forward_" + self.suf)(self.h)
        return self

    def backward(self, dimg):
        """Adjoint of the image for dL/dimage `dimg` (3 x H x W)."""
        d = np.ascontiguousarray(dimg, dtype=self.dtype).reshape(3, self.H, self.W)
        getattr(self.L, "orc_backward_" + self.suf)(self.h, _ptr(d))
        return self

    def gauss_grads_from_raw(self, raw12):
        """Run only the per-Gaussian adjoint on given raw sums (P x 12; rounded
        to float32 first, as the GPU's per-Gaussian kernels round them)."""
        r = np.ascontiguousarray(np.asarray(raw12).astype(np.float32)).reshape(self.P, 12)
        self.L.orc_gauss_grads_from_raw_f32(self.h, _ptr(r))
        return self.grads()

    def get(self, name: str) -> np.ndarray:
        g = getattr(self.L, "orc_get_" + self.suf)
        n = g(self.h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.empty(n, dtype=_DTYPES.get(name, self.dtype))
        if n:
            g(self.h, name.encode(), out.ctypes.data_as(ctypes.c_void_p))
        return out

    def image(self):
        return self.get("out").reshape(3, self.H, self.W)

    def stats(self, accum, accum_abs, max_radii, denom):
        self.L.orc_stats_f32(self.h, _ptr(accum), _ptr(accum_abs), _ptr(max_radii), _ptr(denom))

    def render_simp(self):
        """(accum weights (double), accum counts (int32), max counts (double)) per Gaussian."""
        P = self.P
        w = np.zeros(P, np.float64)
        c = np.zeros(P, np.int32)
        m = np.zeros(P, np.float64)
        self.L.orc_render_simp_f32(self.h, _ptr(w), _ptr(c), _ptr(m))
        return w, c, m

    def render_depth(self):
        """render_depth outputs: points (3,H,W), depth, final T, index (int32) and
        discriminant (H,W each); pixels are zero until written."""
        H, W = self.H, self.W
        pts = np.zeros((3, H, W), np.float32)
        depth = np.zeros((H, W), np.float32)
        alpha = np.zeros((H, W), np.float32)
        gidx = np.zeros((H, W), np.int32)
        disc = np.zeros((H, W), np.float32)
        self.L.orc_render_depth_f32(self.h, _ptr(pts), _ptr(depth), _ptr(alpha), _ptr(gidx),
                                    _ptr(disc))
        return pts, depth, alpha, gidx, disc

    def grads(self) -> dict:
        P = self.P
        return {
            "means3D": self.get("dmean3d").reshape(P, 3),
            "dc": self.get("ddc").reshape(P, 1, 3),
            "shs": self.get("dsh").reshape(P, -1, 3),
            "colors": self.get("dcolor").reshape(P, 3),
            "opacities": self.get("dopac").reshape(P, 1),
            "scales": self.get("dscale").reshape(P, 3),
            "rotations": self.get("drot").reshape(P, 4),
            "cov3D": self.get("dcov3d").reshape(P, 6),
            "mean2D": self.get("dmean2d").reshape(P, 4),
            "conic": self.get("dconic").reshape(P, 4),
        }

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            getattr(self.L, "orc_destroy_" + self.suf)(h)
            self.h = None


def exp2f(x: float) -> float:
    return lib().orc_exp2f(x)


def log2f(x: float) -> float:
    return lib().orc_log2f(x)


def tile_hit(tx, ty, x, y, conic, power, tile_w: int = 16) -> bool:
    return bool(lib().orc_tile_hit_f32(tx, ty, x, y, conic[0], conic[1], conic[2], power,
                                       tile_w))


def num_threads() -> int:
    return lib().orc_num_threads()
