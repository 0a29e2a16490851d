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
                            options=RasterOptions(schedule=schedule, bucket_size=32), headroom=headroom)


def _eager(d, cam, schedule):
    fr = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=d["shs"], options=RasterOptions(schedule=schedule, bucket_size=32)).forward()
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


@pytest.mark.parametrize("tile_w", [16, 32])
def test_replay_matches_oracle(tile_w):
    scene = _scene()
    d = to_dev(scene)
    cams = scenes.orbit_cameras(4, W, H, radius=3.0)
    rend = CapturedRenderer(settings_for(cams[0]), d["means3D"], d["opacities"], d["scales"],
                            d["rotations"], dc=d["dc"], shs=d["shs"],
                            options=RasterOptions(schedule=SCHEDULE_FAST, bucket_size=32,
                                                  tile_w=tile_w), headroom=1.5)
    for k in (1, 3):
        img = rend.render(*_cam_tensors(cams[k])).cpu().numpy()
        orc = Oracle(scene, cams[k], bucket=32, mode=0, tile_w=tile_w).forward()
        np.testing.assert_array_equal(img.view(np.uint32), orc.image().view(np.uint32))


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


# ---------------------------------------------------------------------------
# Captured training steps (forward, L2 loss gradient, backward, statistics)

def _views(k, sizes):
    cams = scenes.orbit_cameras(6, 64, 64, radius=3.0)
    out = []
    for j, (w, h) in enumerate(sizes):
        c = cams[(k + 2 * j) % 6]
        out.append(scenes.look_at_camera(tuple(c.campos), w, h))
    return out


def _eager_step(d, cams, targets, stats):
    from fastergs_amd.multiview import BatchFrame
    fr = BatchFrame([settings_for(c) for c in cams], d["means3D"], d["opacities"], d["scales"],
                    d["rotations"], dc=d["dc"], shs=d["shs"], abs_grad=True).forward()
    scale = torch.cat([torch.full((3 * c.height * c.width,), 2.0 / (3 * c.height * c.width),
                                  device="cuda") for c in cams])
    dimg = (fr.out_color - torch.cat([t.reshape(-1) for t in targets])) * scale
    g = fr.backward(dimg)
    if stats is not None:
        fr.densify_stats(g["mean2D"], *stats)
    torch.cuda.synchronize()
    return fr, g


@pytest.mark.parametrize("sizes", [[(320, 240)], [(200, 150), (320, 240), (160, 96)]],
                         ids=["one_view", "three_views"])
def test_captured_training_step_matches_eager(sizes):
    """Replays of a captured step for changing cameras and targets: images
    and radii bit for bit, gradients and densification statistics to float
    atomics' reassociation, with eager work and host synchronisation between
    replays."""
    from fastergs_amd.graph import CapturedTrainingStep
    from tests.helpers import grad_close
    print(f"\n[graph] captured step, views {sizes}", flush=True)
    d = to_dev(_scene())
    P = d["means3D"].shape[0]
    z = lambda *s: torch.zeros(*s, device="cuda")
    stats = (z(P, 1), z(P, 1), z(P), z(P, 1))
    cams0 = _views(0, sizes)
    step = CapturedTrainingStep([settings_for(c) for c in cams0], d["means3D"], d["opacities"],
                                d["scales"], d["rotations"], dc=d["dc"], shs=d["shs"],
                                xyz_gradient_accum=stats[0], xyz_gradient_accum_abs=stats[1],
                                max_radii=stats[2], denom=stats[3], headroom=1.5)
    ref_stats = tuple(t.clone() for t in stats)
    for k in (0, 1, 3, 0):
        cams = _views(k, sizes)
        step.set_cameras([_cam_tensors(c) for c in cams])
        tg = [torch.from_numpy(scenes.target_image(c.width, c.height, seed=k)).cuda() for c in cams]
        for t, src in zip(step.targets(), tg):
            t.copy_(src)
        g = {n: t.clone() for n, t in step.step().items()}
        img = step.out_color.clone()
        radii = step.radii.clone()
        torch.cuda.synchronize()
        fr, ref = _eager_step(d, cams, tg, ref_stats)  # eager work between replays
        np.testing.assert_array_equal(img.cpu().numpy(), fr.out_color.cpu().numpy())
        np.testing.assert_array_equal(radii.cpu().numpy(), fr.radii.cpu().numpy())
        for n in ("opacities", "dc", "shs", "mean2D", "conic", "means3D", "scales", "rotations"):
            grad_close(g[n].cpu().numpy(), ref[n].cpu().numpy(), rtol=1e-4, name=n)
    assert step.check()
    np.testing.assert_array_equal(stats[3].cpu().numpy(), ref_stats[3].cpu().numpy())
    np.testing.assert_array_equal(stats[2].cpu().numpy(), ref_stats[2].cpu().numpy())
    for a, b in ((stats[0], ref_stats[0]), (stats[1], ref_stats[1])):
        a, b = a.cpu().numpy().ravel(), b.cpu().numpy().ravel()
        assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b)


def test_captured_training_step_capacity():
    """A capacity below the pair count is reported, and a grown capture
    matches the eager step again."""
    from fastergs_amd.graph import CapturedTrainingStep
    d = to_dev(_scene())
    near = scenes.look_at_camera((0.0, 0.0, -2.0), W, H)
    far = scenes.look_at_camera((0.0, 0.0, -6.0), W, H)
    step = CapturedTrainingStep([settings_for(far)], d["means3D"], d["opacities"], d["scales"],
                                d["rotations"], dc=d["dc"], shs=d["shs"], headroom=1.0)
    step.set_cameras([_cam_tensors(near)])
    step.step()
    with pytest.raises(CapacityError):
        step.check()
    step.step()
    assert not step.check(grow=True)
    tg = torch.from_numpy(scenes.target_image(W, H, seed=2)).cuda()
    step.targets()[0].copy_(tg)
    g = {n: t.clone() for n, t in step.step().items()}
    assert step.check()
    fr, ref = _eager_step(d, [near], [tg], None)
    np.testing.assert_array_equal(step.out_color.cpu().numpy(), fr.out_color.cpu().numpy())
    a, b = g["opacities"].cpu().numpy().ravel(), ref["opacities"].cpu().numpy().ravel()
    assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b)
