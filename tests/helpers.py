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


CONIC_FIELDS = (5, 6, 7)  # fgs_math.h RAW_KXX .. RAW_KYY


def raw_close(got, ref, nf: int = 11, name: str = "raw"):
    """Raw per-Gaussian sums (n x >= nf). The conic sums are added
    order-independently (fp64) from the same float wave totals as the
    oracle's: equal to fp64 reassociation (1e-10 relative). The others are
    added with float atomics across waves: equal to float reassociation,
    whose error scales with the summed terms rather than with a cancelling
    sum (1e-4 relative, with the rms floor of grad_close)."""
    got = np.asarray(got, dtype=np.float64).reshape(-1, np.shape(got)[-1])
    ref = np.asarray(ref, dtype=np.float64).reshape(-1, np.shape(ref)[-1])
    for f in range(nf):
        grad_close(got[:, f], ref[:, f], rtol=1e-10 if f in CONIC_FIELDS else 1e-4,
                   name=f"{name}[{f}]")


WELL_CONDITIONED = ("opacities", "dc", "shs", "colors")
GEOMETRIC = ("means3D", "scales", "rotations", "cov3D")


def rel_l2(got, ref) -> float:
    a = np.asarray(got, dtype=np.float64).ravel()
    b = np.asarray(ref, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-300))


def check_backward(got: dict, orc, dimg, abs_grad: bool = True, rtol: float = 1e-4):
    """Backward parity in three parts.

    1. The rasterizer adjoint's raw per-Gaussian sums (colour, mean2D, conic,
       power, |mean2D|; fgs_math.h RawField order), see raw_close: GPU and
       oracle form the same float wave totals; the conic sums agree to fp64
       reassociation, the others to float reassociation.
    2. The oracle's per-Gaussian adjoint applied to the GPU's own raw sums
       reproduces the GPU's gradients to float rounding (1e-5).
    3. End to end against the oracle, every gradient tensor within `rtol`
       (1e-4) elementwise, |got - ref| <= rtol (|ref| + rms(ref)), and in
       relative L2 norm.
    """
    orc.backward(dimg)
    ref = orc.grads()
    raw_ref = orc.get("raw").reshape(-1, 11)
    raw = got["raw"].detach().cpu().numpy().astype(np.float64)
    nf = 11 if abs_grad else 9
    raw_close(raw, raw_ref, nf)
    if not abs_grad:
        assert (raw[:, 9:] == 0).all()
    mine = orc.gauss_grads_from_raw(raw)
    for k in WELL_CONDITIONED + GEOMETRIC + ("mean2D", "conic"):
        a = got[k].detach().cpu().numpy().reshape(mine[k].shape)
        grad_close(a, mine[k], rtol=1e-5, name=f"{k} (adjoint of the GPU raw sums)")
    for k in WELL_CONDITIONED + GEOMETRIC:
        a = got[k].detach().cpu().numpy()
        grad_close(a, ref[k], rtol=rtol, name=k)
        assert rel_l2(a, ref[k]) <= rtol, (k, rel_l2(a, ref[k]))
    return ref
