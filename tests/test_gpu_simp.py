"""render_simp (importance statistics) on the GPU against the oracle, for
both schedules: per-Gaussian counts of contributing pixels and of pixels
where the Gaussian has the largest weight exactly, weight sums within 1e-5
(float atomics against a double sum). Includes image sizes that are not a
multiple of the tile (edge tiles never stop early) and the rasterizer entry
point."""
import numpy as np
import pytest
import torch

from fastergs_amd import scenes
from fastergs_amd.rasterizer import Frame, GaussianRasterizer, RasterOptions
from oracle.pyoracle import Oracle
from tests.helpers import grad_close, settings_for, to_dev

pytestmark = pytest.mark.gpu

CASES = {
    "uniform_20k_300x200": lambda: (scenes.uniform_scene(20000, seed=3),
                                    scenes.look_at_camera((0.0, 0.2, -4.0), 300, 200)),
    "clusters_40k_410x250": lambda: (scenes.cluster_scene(40000, seed=4),
                                     scenes.look_at_camera((0.4, -0.1, -4.2), 410, 250)),
}


SHAPED = [(c, 16) for c in CASES] + [("clusters_40k_410x250", 8), ("clusters_40k_410x250", 32)]


@pytest.mark.parametrize("case,tile_w", SHAPED, ids=[f"{c}-tile{w}" for c, w in SHAPED])
@pytest.mark.parametrize("schedule", [0, 1], ids=["plain", "fast"])
def test_render_simp_matches_oracle(case, tile_w, schedule):
    print(f"\n[simp] {case} tile_w {tile_w} schedule {schedule}", flush=True)
    scene, cam = CASES[case]()
    d = to_dev(scene)
    fr = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=d["shs"], options=RasterOptions(schedule=schedule, bucket_size=32,
                                                               tile_w=tile_w)).forward()
    w, c, m = fr.render_simp()
    orc = Oracle(scene, cam, bucket=32, mode=0 if schedule else 1, tile_w=tile_w).forward()
    rw, rc, rm = orc.render_simp()
    np.testing.assert_array_equal(c.cpu().numpy(), rc)
    np.testing.assert_array_equal(m.cpu().numpy(), rm.astype(np.float32))
    grad_close(w.cpu().numpy(), rw, rtol=1e-5, name="accum weights")
    assert rc.sum() > 0 and rm.sum() > 0


def test_rasterizer_entry_point():
    scene, cam = CASES["uniform_20k_300x200"]()
    d = to_dev(scene)
    ras = GaussianRasterizer(settings_for(cam))
    color, radii, w, c, m = ras.render_simp(d["means3D"], None, d["opacities"], dc=d["dc"],
                                            shs=d["shs"], scales=d["scales"],
                                            rotations=d["rotations"])
    assert color.shape == (3, cam.height, cam.width) and radii.shape == (scene["means3D"].shape[0],)
    # every pixel whose blend met at least one splat credits exactly one maximum
    assert 0 < m.sum().item() <= cam.width * cam.height
    assert (c[radii == 0] == 0).all() and (w >= 0).all()
