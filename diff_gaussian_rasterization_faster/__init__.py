"""Import name used by the FasterGS training/rendering scripts
(``from diff_gaussian_rasterization_faster import GaussianRasterizationSettings,
GaussianRasterizer``, ``SparseGaussianAdam``); forwards to fastergs_amd.
``_C`` is the PyTorch-ROCm extension module whose ``ops`` the autograd
function calls, as the reference package's ``_C`` (__init__.py:15)."""
from fastergs_amd import _C  # noqa: F401
from fastergs_amd.optim import SparseGaussianAdam  # noqa: F401
from fastergs_amd.rasterizer import (  # noqa: F401
    GaussianRasterizationSettings,
    GaussianRasterizer,
    rasterize_gaussians,
)
