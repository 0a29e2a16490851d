"""Shared test helpers: settings from a scene camera, oracle runs, comparisons."""
from __future__ import annotations

import numpy as np
import torch

from fastergs_amd import scenes


def settings_for(cam: scenes.Camera, device="cuda", bg=(0.0, 0.0, 0.0), sh_degree=3,
                 scale_modifier=1.0, mult=1.0):
    from fastergs_amd.rasterizer import GaussianRasterizationSettings
    return GaussianRasterizationSettings(
        image_height=cam.height, image_width=cam.width, tanfovx=cam.tanfovx,
        tanfovy=cam.tanfovy, bg=torch.tensor(bg, dtype=torch.float32, device=device),
        scale_modifier=scale_modifier, viewmatrix=torch.from_numpy(cam.viewmatrix).to(device),
        projmatrix=torch.from_numpy(cam.projmatrix).to(device), sh_degree=sh_degree,
        campos=torch.from_numpy(cam.campos).to(device), mult=mult, prefiltered=False, debug=False)


def to_dev(scene: dict, device="cuda") -> dict:
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in scene.items()}


def grad_close(got: np.ndarray, ref: np.ndarray, rtol=1e-4, name=""):
    """Gradient tolerance used throughout: |got - ref| <= rtol * (|ref| + rms(ref)),
    elementwise; the rms term keeps near-zero entries from demanding more
    than rtol of the tensor's typical magnitude (float32 atomics reassociate
    sums)."""
    got = np.asarray(got, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    assert got.shape == ref.shape, (name, got.shape, ref.shape)
    scale = np.sqrt(np.mean(ref ** 2)) if ref.size else 0.0
    err = np.abs(got - ref)
    bound = rtol * (np.abs(ref) + scale) + 1e-30
    bad = err > bound
    if bad.any():
        i = int(np.argmax(err / bound))
        raise AssertionError(f"{name}: {int(bad.sum())}/{ref.size} entries off; worst idx {i} "
                             f"got {got[i]:.8e} ref {ref[i]:.8e} (rms {scale:.3e})")
