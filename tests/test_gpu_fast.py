"""GPU parity of the FAST schedule against the oracle's bucketed mode.

Bit-exact: records, per-tile lists (and the 64-bit keys rebuilt from them),
tile ranges, bucket offsets / ranges / owner tiles, per-bucket checkpoints up
to each pixel's cutoff, images, T_final, last contributors. Gradients within
1e-4 (relative). The split threshold of the packed/split tile walk and the
bucket size must not change the lists. Every check runs on 16 x 16 tiles (the
reference's) and, for the heavy-tailed case, on 8 x 32 and 32 x 8 tiles
against the oracle on the same tile shape.
"""
import numpy as np
import pytest
import torch

from fastergs_amd import scenes
from fastergs_amd._lib import SCHEDULE_FAST
from fastergs_amd.rasterizer import Frame, RasterOptions
from oracle.pyoracle import Oracle
from tests.helpers import check_backward, grad_close, settings_for, to_dev

pytestmark = pytest.mark.gpu

CASES = {
    "uniform_30k_400x300": lambda: (scenes.uniform_scene(30000, seed=1),
                                    scenes.look_at_camera((0.0, 0.0, -4.0), 400, 300)),
    "heavy_tailed_60k_640x420": lambda: (scenes.cluster_scene(60000, seed=2),
                                         scenes.look_at_camera((0.3, -0.2, -4.5), 640, 420)),
}


def run_fast(scene, cam, bucket=32, split=64, tile_w=16):
    d = to_dev(scene)
    fr = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=d["shs"],
               options=RasterOptions(schedule=SCHEDULE_FAST, bucket_size=bucket,
                                     split_tiles=split, tile_w=tile_w)).forward()
    torch.cuda.synchronize()
    return fr


PARAMS = [(c, 16) for c in CASES] + [("heavy_tailed_60k_640x420", 8),
                                     ("heavy_tailed_60k_640x420", 32)]


@pytest.fixture(scope="module", params=PARAMS, ids=[f"{c}-tile{w}" for c, w in PARAMS])
def case(request):
    name, tile_w = request.param
    print(f"\n[case] {name} tile_w {tile_w}", flush=True)
    scene, cam = CASES[name]()
    fr = run_fast(scene, cam, tile_w=tile_w)
    orc = Oracle(scene, cam, bucket=32, mode=0, tile_w=tile_w).forward()
    return scene, cam, fr, orc


def rebuilt_keys(fr, v):
    vals = v["sorted_vals"].cpu().numpy().view(np.uint32)
    rng = v["ranges"].cpu().numpy().view(np.uint32).reshape(-1, 2)
    tiles = np.repeat(np.arange(rng.shape[0], dtype=np.uint64), (rng[:, 1] - rng[:, 0]))
    depth = v["splat"].cpu().numpy().reshape(-1, 12)[:, 6].view(np.uint32).astype(np.uint64)
    return (tiles << np.uint64(32)) | depth[vals]


def test_lists_ranges_keys(case):
    scene, cam, fr, orc = case
    assert fr.R == orc.get("vals").size
    np.testing.assert_array_equal(fr.radii.cpu().numpy(), orc.get("radii"))
    v = fr.view_arrays()
    np.testing.assert_array_equal(v["pair_count"].cpu().numpy().view(np.uint32),
                                  orc.get("pair_count"))
    np.testing.assert_array_equal(v["sorted_vals"].cpu().numpy().view(np.uint32), orc.get("vals"))
    np.testing.assert_array_equal(v["ranges"].cpu().numpy().view(np.uint32), orc.get("ranges"))
    np.testing.assert_array_equal(rebuilt_keys(fr, v), orc.get("keys"))


def test_bucket_tables(case):
    scene, cam, fr, orc = case
    v = fr.view_arrays()
    boff = v["bucket_off"].cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(boff, orc.get("bucket_off"))
    nb = int(boff[-1])
    assert nb == orc.get("bucket_tile").size <= fr.NB
    np.testing.assert_array_equal(v["bucket_ranges"].cpu().numpy().view(np.uint32)[:2 * nb],
                                  orc.get("bucket_ranges"))
    np.testing.assert_array_equal(v["bucket_tile"].cpu().numpy().view(np.uint32)[:nb],
                                  orc.get("bucket_tile"))


def test_image_and_checkpoints(case):
    scene, cam, fr, orc = case
    np.testing.assert_array_equal(fr.out_color.cpu().numpy().view(np.uint32),
                                  orc.image().view(np.uint32))
    np.testing.assert_array_equal(fr.T_final.cpu().numpy().ravel().view(np.uint32),
                                  orc.get("tfinal").view(np.uint32))
    contrib = orc.get("contrib")
    np.testing.assert_array_equal(fr.last_contributor.cpu().numpy().ravel(), contrib)
    v = fr.view_arrays()
    np.testing.assert_array_equal(v["max_contrib"].cpu().numpy().view(np.uint32),
                                  orc.get("max_contrib"))
    # checkpoints of every (bucket, pixel) up to the pixel's cutoff
    nb = orc.get("bucket_tile").size
    acc = v["acc_c"].cpu().numpy().reshape(-1, 256, 4)[:nb]
    ot = orc.get("acc_t").reshape(nb, 256)
    oc = orc.get("acc_c").reshape(nb, 256, 3)
    br = orc.get("bucket_ranges").reshape(-1, 2).astype(np.int64)
    bt = orc.get("bucket_tile").astype(np.int64)
    rng = orc.get("ranges").reshape(-1, 2).astype(np.int64)
    gx, tw, th = orc.gx, orc.tile_w, orc.tile_h
    lp = np.arange(256)
    tx, ty = bt % gx, bt // gx
    px = (tx[:, None] * tw + (lp % tw)[None, :])
    py = (ty[:, None] * th + (lp // tw)[None, :])
    inside = (px < orc.W) & (py < orc.H)
    c = np.zeros_like(px)
    c[inside] = contrib.reshape(orc.H, orc.W)[py[inside], px[inside]]
    local_start = (br[:, 0] - rng[bt, 0])[:, None]
    valid = inside & (local_start < c)
    assert valid.any()
    np.testing.assert_array_equal(acc[..., 3][valid].view(np.uint32), ot[valid].view(np.uint32))
    np.testing.assert_array_equal(acc[..., :3][valid].view(np.uint32), oc[valid].view(np.uint32))


@pytest.mark.parametrize("abs_grad", [True, False])
def test_backward_parity(case, abs_grad):
    scene, cam, fr, orc = case
    tgt = scenes.target_image(cam.width, cam.height)
    img = orc.image()
    dimg = (2.0 * (img - tgt) / img.size).astype(np.float32)
    fr.copts.abs_grad = 1 if abs_grad else 0
    got = fr.backward(torch.from_numpy(dimg).cuda())
    torch.cuda.synchronize()
    check_backward(got, orc, dimg, abs_grad)


@pytest.mark.parametrize("split", [1, 16, 4096])
def test_split_threshold_does_not_change_lists(split):
    scene = scenes.cluster_scene(20000, seed=9)
    cam = scenes.look_at_camera((0.0, 0.3, -4.0), 480, 320)
    ref = run_fast(scene, cam, split=64).view_arrays()
    got = run_fast(scene, cam, split=split).view_arrays()
    for k in ("pair_count", "sorted_vals", "ranges", "bucket_off"):
        assert torch.equal(got[k], ref[k]), k


@pytest.mark.parametrize("bucket", [16, 64, 128, 256])
def test_bucket_sizes(bucket):
    scene = scenes.uniform_scene(15000, seed=4)
    cam = scenes.look_at_camera((0.2, 0.0, -4.0), 300, 220)
    fr = run_fast(scene, cam, bucket=bucket)
    orc = Oracle(scene, cam, bucket=bucket, mode=0).forward()
    np.testing.assert_array_equal(fr.out_color.cpu().numpy().view(np.uint32),
                                  orc.image().view(np.uint32))
    np.testing.assert_array_equal(fr.last_contributor.cpu().numpy().ravel(), orc.get("contrib"))
    img = orc.image()
    dimg = (img - 0.5).astype(np.float32)
    got = fr.backward(torch.from_numpy(dimg).cuda())
    check_backward(got, orc, dimg)
