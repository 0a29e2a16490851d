"""gfx950-native differentiable Gaussian-splatting rasterizer.

Drop-in replacement for FasterGS's ``diff_gaussian_rasterization_faster``
(see ``fastergs_amd.rasterizer``); compute runs in libfastergs.so.
"""
from .rasterizer import (  # noqa: F401
    Frame,
    GaussianRasterizationSettings,
    GaussianRasterizer,
    RasterOptions,
    get_default_options,
    rasterize_gaussians,
    set_default_options,
)
from ._lib import SCHEDULE_FAST, SCHEDULE_PLAIN, FasterGSError  # noqa: F401
from .optim import SparseGaussianAdam  # noqa: F401
