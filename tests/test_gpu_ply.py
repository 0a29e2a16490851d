"""Real-scene replay path on the GPU: a model written as a 3DGS .ply of raw
parameters (logit opacity, log scales, unnormalised quaternions) and its
cameras.json are loaded back, rendered through the drop-in API, and compared
with the direct render of the same activated attributes (bit for bit) and
with the oracle on the loaded cameras; bench.py's replay of the files runs."""
import argparse

import numpy as np
import pytest
import torch

from fastergs_amd import plyio, scenes
from fastergs_amd.rasterizer import GaussianRasterizer, RasterOptions, set_default_options
from oracle.pyoracle import Oracle
from tests.helpers import settings_for, to_dev

pytestmark = pytest.mark.gpu


def _raw_scene(n, seed):
    rng = np.random.default_rng(seed)
    sc = scenes.cluster_scene(n, seed=seed, clusters=32, radius=1.5)
    raw = {
        "means3D": sc["means3D"],
        "opacities": rng.normal(0.0, 1.5, (n, 1)).astype(np.float32),       # logits
        "scales": np.log(sc["scales"]).astype(np.float32),                   # log scales
        "rotations": (sc["rotations"] * rng.uniform(0.5, 2.0, (n, 1))).astype(np.float32),
        "dc": sc["dc"], "shs": sc["shs"],
    }
    return raw


def _activate(raw):
    """The trainers' getters (sigmoid, exp, normalise) in float64, as load_ply's."""
    a = dict(raw)
    a["opacities"] = (1.0 / (1.0 + np.exp(-raw["opacities"].astype(np.float64)))).astype(np.float32)
    a["scales"] = np.exp(raw["scales"].astype(np.float64)).astype(np.float32)
    q = raw["rotations"].astype(np.float64)
    a["rotations"] = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    return a


def _render(scene, cam):
    d = to_dev(scene)
    with torch.no_grad():
        img, radii = GaussianRasterizer(settings_for(cam))(
            means3D=d["means3D"], opacities=d["opacities"], dc=d["dc"], shs=d["shs"],
            scales=d["scales"], rotations=d["rotations"])
    torch.cuda.synchronize()
    return img.cpu().numpy(), radii.cpu().numpy()


def test_ply_and_cameras_round_trip(tmp_path):
    set_default_options(schedule=1, bucket_size=32)
    try:
        raw = _raw_scene(30000, seed=3)
        ply = str(tmp_path / "point_cloud.ply")
        plyio.save_ply(ply, raw, activated=False)
        loaded = plyio.load_ply(ply)
        assert loaded.pop("sh_degree") == 3
        direct = _activate(raw)
        for k in direct:
            np.testing.assert_array_equal(loaded[k], direct[k], err_msg=k)
        cams = [scenes.look_at_camera(tuple(c.campos), 400, 300) for c in
                scenes.orbit_cameras(3, 400, 300, radius=3.0)]
        cj = str(tmp_path / "cameras.json")
        plyio.save_cameras_json(cj, cams)
        back = plyio.load_cameras_json(cj)
        assert len(back) == 3
        for c in back:
            img, radii = _render(loaded, c)
            ref, ref_r = _render(direct, c)
            np.testing.assert_array_equal(img.view(np.uint32), ref.view(np.uint32))
            orc = Oracle(loaded, c, bucket=32, mode=0).forward()
            np.testing.assert_array_equal(img.view(np.uint32), orc.image().view(np.uint32))
            np.testing.assert_array_equal(radii, orc.get("radii"))
            assert (radii > 0).sum() > 1000
        # the benchmark's replay of these files (forward FPS and fwd+bwd per view)
        import bench
        rep = bench.scene_replay(argparse.Namespace(ply=ply, cameras=cj), max_views=3)
        assert rep["views"] == 3 and rep["gaussians"] == 30000 and rep["fps"] > 0
        assert rep["fwd_bwd_ms"] > 0
    finally:
        set_default_options(**vars(RasterOptions()))
