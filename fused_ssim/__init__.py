"""Import name used by the FasterGS training loops (``from fused_ssim import
fused_ssim``); forwards to fastergs_amd.ssim."""
from fastergs_amd.ssim import FusedSSIMMap, allowed_padding, fused_ssim  # noqa: F401
