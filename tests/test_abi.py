"""CPU tests of the C-ABI boundary: libfastergs.so loads, exports every
function include/fastergs.h declares, the Python binding's signature table
matches the header, and host-side argument checking works without a GPU."""
import ctypes
import os
import re

import pytest

from fastergs_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "fastergs.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fgs_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_and_library_exports():
    names = header_functions()
    assert len(names) >= 15
    L = _lib.lib()
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/fastergs.h but not exported"
    assert set(names) == set(_lib.SIGNATURES), "ctypes signature table out of sync with header"


def test_workspace_queries_and_build_info():
    L = _lib.lib()
    opt = _lib.Options(0, 32, 64, 1)
    g = L.fgs_gauss_ws_size(1000, 640, 480, ctypes.byref(opt))
    assert g > 1000 * (32 + 48)
    p = L.fgs_pair_ws_size(1000, 640, 480, 50000, 2000, ctypes.byref(opt))
    assert p > 50000 * 24
    fast = _lib.Options(1, 32, 64, 1)
    pf = L.fgs_pair_ws_size(1000, 640, 480, 50000, 2000, ctypes.byref(fast))
    assert pf > 2000 * 256 * 16  # bucket checkpoints
    assert L.fgs_sort_ws_size(1 << 20) > 256 * 4 * 256
    assert L.fgs_grad_ws_size(10) >= 10 * 48
    assert b"gfx950" in L.fgs_build_info()


def test_argument_errors_without_gpu():
    L = _lib.lib()
    bad = _lib.Options(0, 33, 64, 1)  # bucket size not allowed
    cam = _lib.Camera(64, 64, 0.5, 0.5, 1.0, 1.0, 3, None, None, None, None)
    gs = _lib.Gaussians(0, 15, None, None, None, None, None, None, None, None)
    rc = L.fgs_preprocess(ctypes.byref(cam), ctypes.byref(gs), ctypes.byref(bad), None, 0, None,
                          None)
    assert rc == -1 and b"bucket_size" in L.fgs_last_error()
    ok = _lib.Options(0, 32, 64, 1)
    rc = L.fgs_preprocess(ctypes.byref(cam), ctypes.byref(gs), ctypes.byref(ok), None, 0, None,
                          None)
    assert rc == -1 and b"device pointers" in L.fgs_last_error()
    with pytest.raises(_lib.FasterGSError):
        _lib.check(-1, "probe")


def test_drop_in_import_names():
    import diff_gaussian_rasterization_faster as d
    for n in ("GaussianRasterizationSettings", "GaussianRasterizer", "rasterize_gaussians"):
        assert hasattr(d, n)
    fields = d.GaussianRasterizationSettings._fields
    assert fields == ("image_height", "image_width", "tanfovx", "tanfovy", "bg",
                      "scale_modifier", "viewmatrix", "projmatrix", "sh_degree", "campos", "mult",
                      "prefiltered", "debug")


def test_noalloc_sort_rejects_sizes_beyond_its_scratch():
    """Host-side check, before any device work: more than 2^24 pairs is a
    capacity error (the pointers are never touched)."""
    L = _lib.lib()
    fake = ctypes.c_void_p(0x1000)
    rc = L.fgs_sort_pairs_noalloc(fake, fake, fake, fake, (1 << 24) + 1, 0, 45, None)
    assert rc == -3  # FGS_ERR_CAPACITY
    assert b"2^24" in L.fgs_last_error()
    assert L.fgs_sort_pairs_noalloc(None, None, None, None, -1, 0, 45, None) == -1


def test_forward_only_pair_workspace_is_smaller():
    """A forward-only frame's pair workspace holds no bucket checkpoints
    (NB x 256 x 16 B) and no per-pixel colours (H x W x 16 B)."""
    L = _lib.lib()
    full = _lib.Options(1, 32, 64, 1)
    fo = _lib.Options(1, 32, 64, 1)
    fo.forward_only = 1
    R, NB, W, H = 500000, 20000, 640, 480
    a = L.fgs_pair_ws_size(1000, W, H, R, NB, ctypes.byref(full))
    b = L.fgs_pair_ws_size(1000, W, H, R, NB, ctypes.byref(fo))
    assert a - b >= NB * 256 * 16 + W * H * 16


def test_torch_extension_loads_and_shares_the_library():
    """The PyTorch-ROCm extension imports (without a GPU), exposes its ops,
    and resolves libfastergs.so to the same in-tree file ctypes loads."""
    import diff_gaussian_rasterization_faster as d
    from fastergs_amd import _C
    assert d._C is _C
    assert {"forward", "backward", "densify_stats"} <= set(dir(_C.ops))
    assert _C.build_info() == _lib.lib().fgs_build_info().decode()
    maps = open("/proc/self/maps").read()
    libs = {line.split()[-1] for line in maps.splitlines() if line.endswith("libfastergs.so")}
    assert len(libs) == 1, libs
