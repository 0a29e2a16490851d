"""GPU tests of prefix lists (RasterOptions.list_prefix_div / fgs_options
list_prefix): a FAST frame first built from the per-tile lists of its
nearest Gaussians equals the whole-list frame and the oracle bit for bit
(image, T_final, last contributors; gradients within 1e-4 like every other
gradient test), whether the prefixes sufficed or the device built the whole
lists and composited again. Which of the two happened is read from the
frame's counts words: [7] the prefix pair count, [8] the redo flag."""
import ctypes

import numpy as np
import pytest
import torch

from fastergs_amd import rasterizer, scenes
from fastergs_amd._lib import FrameView, check, lib
from fastergs_amd.rasterizer import Frame, GaussianRasterizer, RasterOptions
from oracle.pyoracle import Oracle
from tests.helpers import grad_close, settings_for, to_dev

pytestmark = pytest.mark.gpu

COUNT_PREFIX_N, COUNT_REDO = 7, 8


def _opaque():
    """Opaque clusters whose every tile saturates within the nearest 1.3 %
    (bucket 32) / 3.3 % (bucket 128) of the Gaussians at 640 x 420."""
    return (scenes.cluster_scene(60000, 0, opacity_mu=3.0, big_frac=0.1),
            scenes.look_at_camera((0.0, 0.0, -4.0), 640, 420))


def _sparse():
    """Small uniform splats: most tiles never saturate, so prefixes cannot finish them."""
    return scenes.uniform_scene(20000, seed=3), scenes.look_at_camera((0.2, 0.1, -4.0), 320, 240)


def _counts(fr):
    v = FrameView()
    check(lib().fgs_frame_view_get(ctypes.c_void_p(fr.gws.data_ptr()), None, fr.P, fr.W, fr.H, 0,
                                   0, ctypes.byref(fr.copts), ctypes.byref(v)), "view")
    off = v.counts - fr.gws.data_ptr()
    return fr.gws[off:off + 64].cpu().numpy().view(np.uint32)


def _frame(scene, cam, div, bucket, forward_only=False):
    rasterizer._prefix_policies.clear()
    d = to_dev(scene)
    fr = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=d["shs"], forward_only=forward_only,
               options=RasterOptions(bucket_size=bucket, list_prefix_div=div))
    return fr.forward()


def _check_outputs(fr, orc):
    for got, name in ((fr.out_color, "out"), (fr.T_final, "tfinal"),
                      (fr.last_contributor, "contrib")):
        ref = orc.get(name).reshape(-1)
        got = got.detach().cpu().numpy().reshape(-1)
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32), err_msg=name)


def _check_grads(fr, orc, cam):
    dimg = torch.from_numpy(scenes.target_image(cam.width, cam.height, seed=2)).cuda()
    g = fr.backward(dimg)
    orc.backward(dimg.cpu().numpy())
    ref = orc.grads()
    for k in ("opacities", "dc", "shs", "means3D", "scales", "rotations"):
        grad_close(g[k].cpu().numpy(), ref[k], name=k)


@pytest.mark.parametrize("bucket", [32, 128])
def test_prefix_lists_suffice(bucket):
    scene, cam = _opaque()
    fr = _frame(scene, cam, 16, bucket)
    c = _counts(fr)
    assert 0 < c[COUNT_PREFIX_N] < c[0] // 2, (c[COUNT_PREFIX_N], c[0])
    assert c[COUNT_REDO] == 0
    orc = Oracle(scene, cam, bucket=bucket, mode=0).forward()
    _check_outputs(fr, orc)
    _check_grads(fr, orc, cam)


@pytest.mark.parametrize("case", ["sparse", "short_prefix"])
def test_redo_builds_whole_lists(case):
    """Prefixes that leave pixels open: compositing raises the redo flag and
    the whole lists are built and composited again, all on the device."""
    if case == "sparse":
        (scene, cam), div = _sparse(), 16
    else:
        (scene, cam), div = _opaque(), 6000  # the 10 nearest Gaussians
    fr = _frame(scene, cam, div, 32)
    c = _counts(fr)
    assert c[COUNT_REDO] == 1 and c[COUNT_PREFIX_N] > 0
    orc = Oracle(scene, cam, bucket=32, mode=0).forward()
    _check_outputs(fr, orc)
    _check_grads(fr, orc, cam)


def test_complete_lists_for_list_passes():
    """A prefix frame's lists are rebuilt whole before inspection and the
    list passes; its outputs keep their values and the backward still
    reads consistent checkpoints."""
    scene, cam = _opaque()
    fr = _frame(scene, cam, 16, 32)
    before = [t.clone() for t in (fr.out_color, fr.T_final, fr.last_contributor)]
    v = fr.view_arrays()
    orc = Oracle(scene, cam, bucket=32, mode=0).forward()
    np.testing.assert_array_equal(v["ranges"].cpu().numpy().view(np.uint32), orc.get("ranges"))
    np.testing.assert_array_equal(v["sorted_vals"].cpu().numpy().view(np.uint32), orc.get("vals"))
    for a, b in zip(before, (fr.out_color, fr.T_final, fr.last_contributor)):
        assert torch.equal(a, b)
    assert _counts(fr)[COUNT_PREFIX_N] == 0
    _check_grads(fr, orc, cam)


def test_forward_only_prefix():
    scene, cam = _opaque()
    fr = _frame(scene, cam, 16, 128, forward_only=True)
    assert _counts(fr)[COUNT_REDO] == 0
    _check_outputs(fr, Oracle(scene, cam, bucket=128, mode=0).forward())


def test_status_word_turns_prefixes_off():
    """The drop-in API reads each prefix frame's status word once the frame
    has completed: after a frame whose prefixes did not suffice, frames of
    that shape render whole lists; images match the oracle throughout."""
    rasterizer._prefix_policies.clear()
    scene, cam = _sparse()
    d = to_dev(scene)
    ras = GaussianRasterizer(settings_for(cam))
    orc = Oracle(scene, cam, bucket=128, mode=0).forward()
    ref = orc.image().reshape(-1).view(np.uint32)
    used = []
    for _ in range(3):
        with torch.no_grad():
            img, _ = ras(means3D=d["means3D"], opacities=d["opacities"], dc=d["dc"],
                         shs=d["shs"], scales=d["scales"], rotations=d["rotations"])
        torch.cuda.synchronize()
        np.testing.assert_array_equal(img.cpu().numpy().reshape(-1).view(np.uint32), ref)
        (pol,) = rasterizer._prefix_policies.values()
        used.append(pol.off == 0)
    # frame 1 tried prefixes; its status (read at frame 2) turned them off
    assert used == [True, False, False], used
    (pol,) = rasterizer._prefix_policies.values()
    assert pol.off == pol.RETRY - 2
    rasterizer._prefix_policies.clear()


def test_captured_frames_with_prefix_lists():
    """Graph-captured frames take a prefix size fixed at capture (whole lists
    when the capture's warm-up frame needed them): replays match the oracle
    whether the prefixes sufficed or a replay for another camera built the
    whole lists, which prefix_redone() reports."""
    from fastergs_amd.graph import CapturedRenderer

    def renderer(scene, cam):
        d = to_dev(scene)
        return CapturedRenderer(settings_for(cam), d["means3D"], d["opacities"], d["scales"],
                                d["rotations"], dc=d["dc"], shs=d["shs"],
                                options=RasterOptions(bucket_size=32, list_prefix_div=16),
                                headroom=1.5)

    def cam_tensors(c):
        return (torch.from_numpy(c.viewmatrix).cuda(), torch.from_numpy(c.projmatrix).cuda(),
                torch.from_numpy(c.campos).cuda())

    (scene, cam), (sparse, scam) = _opaque(), _sparse()
    rend = renderer(scene, cam)
    assert rend.copts.list_prefix > 0
    far = scenes.look_at_camera((0.0, 0.0, -12.0), cam.width, cam.height)
    for c, redone in ((cam, False), (far, True), (cam, False)):
        img = rend.render(*cam_tensors(c)).cpu().numpy()
        orc = Oracle(scene, c, bucket=32, mode=0).forward()
        np.testing.assert_array_equal(img.view(np.uint32), orc.image().view(np.uint32))
        assert rend.prefix_redone() is redone
    rend = renderer(sparse, scam)
    assert rend.copts.list_prefix == 0  # the warm-up frame needed the whole lists
    img = rend.render().cpu().numpy()
    orc = Oracle(sparse, scam, bucket=32, mode=0).forward()
    np.testing.assert_array_equal(img.view(np.uint32), orc.image().view(np.uint32))


def test_captured_training_step_with_prefix_lists():
    """One-view captured training step on prefix lists against the eager
    whole-list step: image bit for bit, gradients within 1e-4."""
    from fastergs_amd.graph import CapturedTrainingStep
    from fastergs_amd.multiview import BatchFrame
    scene, cam = _opaque()
    d = to_dev(scene)
    step = CapturedTrainingStep([settings_for(cam)], d["means3D"], d["opacities"], d["scales"],
                                d["rotations"], dc=d["dc"], shs=d["shs"],
                                options=RasterOptions(bucket_size=32, list_prefix_div=16),
                                headroom=1.5)
    tg = torch.from_numpy(scenes.target_image(cam.width, cam.height, seed=4)).cuda()
    step.targets()[0].copy_(tg)
    g = {n: t.clone() for n, t in step.step().items()}
    img = step.out_color.clone()
    assert not step.prefix_redone()
    fr = BatchFrame([settings_for(cam)], d["means3D"], d["opacities"], d["scales"],
                    d["rotations"], dc=d["dc"], shs=d["shs"],
                    options=RasterOptions(bucket_size=32, list_prefix_div=0)).forward()
    n = 3 * cam.height * cam.width
    ref = fr.backward((fr.out_color - tg.reshape(-1)) * (2.0 / n))
    np.testing.assert_array_equal(img.cpu().numpy(), fr.out_color.cpu().numpy())
    for k in ("opacities", "dc", "shs", "means3D", "scales", "rotations"):
        grad_close(g[k].cpu().numpy(), ref[k].cpu().numpy(), name=k)


def test_batch_prefix_lists():
    """A batch's depth order interleaves its views' records and keeps each
    view's own order, so its prefix is a prefix of every view's order: the
    batch rendered from prefix lists equals the whole-list batch (images,
    T_final, contributors bit for bit, gradients within 1e-4). The prefix
    is the batch's nearest records, not a share of each view's."""
    from fastergs_amd.multiview import BatchFrame
    scene, cam = _opaque()
    d = to_dev(scene)
    # views at one distance from the scene: the batch's nearest records
    # cover every view (views at unequal distances may need the redo)
    cams = [cam, scenes.look_at_camera((0.3, 0.2, -4.0), 480, 320),
            scenes.look_at_camera((-0.2, 0.0, -4.0), 320, 240)]
    sts = [settings_for(c) for c in cams]
    rasterizer._prefix_policies.clear()

    def batch(div):
        return BatchFrame(sts, d["means3D"], d["opacities"], d["scales"], d["rotations"],
                          dc=d["dc"], shs=d["shs"],
                          options=RasterOptions(bucket_size=32, list_prefix_div=div)).forward()

    a, b = batch(16), batch(0)
    v = FrameView()
    check(lib().fgs_batch_frame_view_get(ctypes.c_void_p(a.gws.data_ptr()), None, a.P,
                                         ctypes.byref(a.batch), 0, 0, ctypes.byref(a.copts),
                                         ctypes.byref(v)), "view")
    off = v.counts - a.gws.data_ptr()
    c = a.gws[off:off + 64].cpu().numpy().view(np.uint32)
    assert 0 < c[COUNT_PREFIX_N] < c[0] and c[COUNT_REDO] == 0, (c[COUNT_PREFIX_N], c[0])
    for x, y in ((a.out_color, b.out_color), (a.T_final, b.T_final),
                 (a.last_contributor, b.last_contributor)):
        assert torch.equal(x, y)
    dimg = torch.randn_like(a.out_color) * 1e-3
    ga, gb = a.backward(dimg), b.backward(dimg)
    for k in ("opacities", "dc", "shs", "means3D", "scales", "rotations"):
        grad_close(ga[k].cpu().numpy(), gb[k].cpu().numpy(), name=k)
    rasterizer._prefix_policies.clear()
