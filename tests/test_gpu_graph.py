"""Graph-captured forward: replays against eager frames and the oracle.

One frame (preprocess, emission, sort, ranges, buckets, compositing) is
captured once and replayed for several cameras of an orbit; every replay is
compared bit for bit with an eager Frame of the same camera and with the
oracle's image, with eager work and host synchronisations between replays.
A deliberately small pair capacity must be reported (never written past)
and a grown capture must again match the eager frame.
"""
import numpy as np
import pytest
import torch

from fastergs_amd import scenes
from fastergs_amd._lib import SCHEDULE_FAST, SCHEDULE_PLAIN
from fastergs_amd.graph import CapacityError, CapturedRenderer
from fastergs_amd.rasterizer import Frame, RasterOptions
from oracle.pyoracle import Oracle
from tests.helpers import settings_for, to_dev

pytestmark = pytest.mark.gpu

W, H = 320, 240


def _scene():
    return scenes.cluster_scene(20000, seed=5, clusters=16, radius=1.5)


def _renderer(d, cam, schedule, headroom=1.5):
    return CapturedRenderer(settings_for(cam), d["means3D"], d["opacities"], d["scales"],
                            d["rotations"], dc=d["dc"], shs=d["shs"],
                            options=RasterOptions(schedule=schedule), headroom=headroom)


def _eager(d, cam, schedule):
    fr = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=d["shs"], options=RasterOptions(schedule=schedule)).forward()
    torch.cuda.synchronize()
    return fr


def _cam_tensors(cam):
    return (torch.from_numpy(cam.viewmatrix).cuda(), torch.from_numpy(cam.projmatrix).cuda(),
            torch.from_numpy(cam.campos).cuda())


@pytest.mark.parametrize("schedule", [SCHEDULE_FAST, SCHEDULE_PLAIN], ids=["fast", "plain"])
def test_replays_match_eager(schedule):
    print(f"\n[graph] replays, schedule {schedule}", flush=True)
    scene = _scene()
    d = to_dev(scene)
    cams = scenes.orbit_cameras(6, W, H, radius=3.0)
    rend = _renderer(d, cams[0], schedule)
    for k in [0, 1, 2, 3, 4, 5, 2, 0]:
        img = rend.render(*_cam_tensors(cams[k])).clone()
        tf = rend.T_final.clone()
        lc = rend.last_contributor.clone()
        torch.cuda.synchronize()
        ref = _eager(d, cams[k], schedule)  # eager work between replays
        np.testing.assert_array_equal(img.cpu().numpy(), ref.out_color.cpu().numpy())
        np.testing.assert_array_equal(tf.cpu().numpy(), ref.T_final.cpu().numpy())
        np.testing.assert_array_equal(lc.cpu().numpy(), ref.last_contributor.cpu().numpy())
        np.testing.assert_array_equal(rend.radii.cpu().numpy(), ref.radii.cpu().numpy())
    assert rend.check()


def test_replay_matches_oracle():
    scene = _scene()
    d = to_dev(scene)
    cams = scenes.orbit_cameras(4, W, H, radius=3.0)
    rend = _renderer(d, cams[0], SCHEDULE_FAST)
    for k in (1, 3):
        img = rend.render(*_cam_tensors(cams[k])).cpu().numpy()
        orc = Oracle(scene, cams[k], bucket=32, mode=0).forward()
        np.testing.assert_allclose(img, orc.image(), atol=1e-5, rtol=0)


def test_capacity_overflow_reported_and_grown():
    print("\n[graph] capacity overflow", flush=True)
    scene = _scene()
    d = to_dev(scene)
    cam = scenes.orbit_cameras(1, W, H, radius=3.0)[0]
    rend = _renderer(d, cam, SCHEDULE_FAST)
    R, NB = rend.capacity()
    # re-capture with half of the pairs this camera needs
    rend._capture(int(R / rend.headroom) // 2, 0)
    rend.render()
    torch.cuda.synchronize()
    with pytest.raises(CapacityError):
        rend.check()
    rend.render()
    assert rend.check(grow=True) is False
    img = rend.render().clone()
    assert rend.check()
    ref = _eager(d, cam, SCHEDULE_FAST)
    np.testing.assert_array_equal(img.cpu().numpy(), ref.out_color.cpu().numpy())
