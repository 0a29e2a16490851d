"""render_depth (per-pixel depth of the dominant splat) on the GPU against the
oracle, for both schedules: index, points, depth, final T and discriminant
bit-exact (the kernel and the oracle share the ray, the inverse and the
ellipsoid arithmetic, compiled without contraction). Includes image sizes
that are not a multiple of the tile, an empty frame and the rasterizer entry
point."""
import numpy as np
import pytest
import torch

from fastergs_amd import scenes
from fastergs_amd.rasterizer import Frame, GaussianRasterizer, RasterOptions
from oracle.pyoracle import Oracle
from tests.helpers import settings_for, to_dev

pytestmark = pytest.mark.gpu

CASES = {
    "uniform_20k_300x200": lambda: (scenes.uniform_scene(20000, seed=3),
                                    scenes.look_at_camera((0.0, 0.2, -4.0), 300, 200)),
    "clusters_40k_410x250": lambda: (scenes.cluster_scene(40000, seed=4),
                                     scenes.look_at_camera((0.4, -0.1, -4.2), 410, 250)),
}


def _frame(scene, cam, schedule, tile_w=16):
    d = to_dev(scene)
    return Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
                 dc=d["dc"], shs=d["shs"], options=RasterOptions(schedule=schedule, bucket_size=32,
                                                                 tile_w=tile_w)).forward()


SHAPED = [(c, 16) for c in CASES] + [("clusters_40k_410x250", 8), ("clusters_40k_410x250", 32)]


@pytest.mark.parametrize("case,tile_w", SHAPED, ids=[f"{c}-tile{w}" for c, w in SHAPED])
@pytest.mark.parametrize("schedule", [0, 1], ids=["plain", "fast"])
def test_render_depth_matches_oracle(case, tile_w, schedule):
    print(f"\n[depth] {case} tile_w {tile_w} schedule {schedule}", flush=True)
    scene, cam = CASES[case]()
    res = _frame(scene, cam, schedule, tile_w).render_depth()
    orc = Oracle(scene, cam, bucket=32, mode=0 if schedule else 1, tile_w=tile_w).forward()
    pts, depth, alpha, gidx, disc = orc.render_depth()
    np.testing.assert_array_equal(res["gidx"][0].cpu().numpy(), gidx)
    np.testing.assert_array_equal(res["accum_alpha"][0].cpu().numpy(), alpha)
    np.testing.assert_array_equal(res["rendered_depth"][0].cpu().numpy(), depth)
    np.testing.assert_array_equal(res["discriminants"][0].cpu().numpy(), disc)
    np.testing.assert_array_equal(res["out_pts"].cpu().numpy(), pts)
    assert (gidx != 0).sum() > cam.width * cam.height // 10
    assert (depth >= 0).all()


def test_empty_frame():
    scene, cam = CASES["uniform_20k_300x200"]()
    scene["opacities"][:] = 0.001          # every alpha below 1/255
    res = _frame(scene, cam, 1).render_depth()
    assert (res["accum_alpha"] == 1.0).all() and (res["gidx"] == 0).all()
    assert (res["rendered_depth"] == 0).all() and (res["out_pts"] == 0).all()


def test_rasterizer_entry_point():
    scene, cam = CASES["uniform_20k_300x200"]()
    d = to_dev(scene)
    ras = GaussianRasterizer(settings_for(cam))
    res = ras.render_depth(d["means3D"], None, d["opacities"], dc=d["dc"], shs=d["shs"],
                           scales=d["scales"], rotations=d["rotations"])
    H, W = cam.height, cam.width
    assert res["render"].shape == (3, H, W) and res["out_pts"].shape == (3, H, W)
    for k in ("rendered_depth", "accum_alpha", "gidx", "discriminants"):
        assert res[k].shape == (1, H, W), k
    assert res["gidx"].dtype == torch.int32
    # the forward's final T is reached only where the walk did not stop early
    tf = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=d["shs"]).forward().T_final.reshape(1, H, W)
    assert (res["accum_alpha"] >= tf - 1e-6).all()
    with pytest.raises(ValueError):
        ras.render_depth(d["means3D"], None, d["opacities"], dc=d["dc"], shs=d["shs"],
                         scales=None, rotations=d["rotations"])
