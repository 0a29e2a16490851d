"""Batches of views (one launch per stage) against single-view frames.

For a batch of cameras with different image sizes:
* every view's image, T_final, last contributors and radii are bit-identical
  to a single-view FAST frame of that camera, and its per-tile lists are
  the single-view lists (record indices v * P + i mapped back to i);
* the screen-space gradients of each view match the single-view ones and
  the parameter gradients match the sum of the single-view gradients
  (1e-4 relative, float atomics reassociate sums);
* the densification statistics equal consecutive single-view updates;
* the autograd entry point returns per-view images and the same gradients.
"""
import numpy as np
import pytest
import torch

from fastergs_amd import scenes
from fastergs_amd._lib import SCHEDULE_FAST
from fastergs_amd.multiview import BatchFrame, rasterize_views
from fastergs_amd.rasterizer import Frame, RasterOptions
from tests.helpers import grad_close, raw_close, settings_for, to_dev

pytestmark = pytest.mark.gpu

SIZES = ((320, 240), (200, 150), (416, 300), (256, 256))


@pytest.fixture(scope="module")
def batch_case():
    scene = scenes.cluster_scene(30000, seed=21, clusters=16, radius=1.5)
    cams = scenes.sphere_cameras(len(SIZES), seed=4, rmin=3.0, rmax=4.5, sizes=SIZES)
    d = to_dev(scene)
    opts = RasterOptions(schedule=SCHEDULE_FAST, bucket_size=32)
    bf = BatchFrame([settings_for(c) for c in cams], d["means3D"], d["opacities"], d["scales"],
                    d["rotations"], dc=d["dc"], shs=d["shs"], options=opts).forward()
    singles = [Frame(settings_for(c), d["means3D"], d["opacities"], d["scales"], d["rotations"],
                     dc=d["dc"], shs=d["shs"], options=opts).forward() for c in cams]
    torch.cuda.synchronize()
    return scene, cams, d, bf, singles


def test_forward_matches_single_views(batch_case):
    print("\n[multiview] forward", flush=True)
    scene, cams, d, bf, singles = batch_case
    imgs = bf.images()
    tf = bf.per_view(bf.T_final)
    lc = bf.per_view(bf.last_contributor)
    assert bf.R == sum(s.R for s in singles)
    for v, s in enumerate(singles):
        np.testing.assert_array_equal(imgs[v].cpu().numpy(), s.out_color.cpu().numpy())
        np.testing.assert_array_equal(tf[v].cpu().numpy(), s.T_final[..., 0].cpu().numpy())
        np.testing.assert_array_equal(lc[v].cpu().numpy(), s.last_contributor[..., 0].cpu().numpy())
        np.testing.assert_array_equal(bf.radii[v].cpu().numpy(), s.radii.cpu().numpy())


def test_queue_compositing_matches(batch_case):
    """The batch composited by the persistent work-queue grid (work items of
    all views taken from one counter) gives the same images, T_final, last
    contributors and raw sums as one workgroup per tile, twice in a row."""
    print("\n[multiview] queue", flush=True)
    scene, cams, d, bf, singles = batch_case
    g = torch.cat([torch.from_numpy(scenes.target_image(c.width, c.height, seed=v)).cuda()
                   .reshape(-1) for v, c in enumerate(cams)])
    ref = bf.backward(g)
    for rep in range(2):
        opts = RasterOptions(schedule=SCHEDULE_FAST, bucket_size=32, composite_queue=True)
        q = BatchFrame([settings_for(c) for c in cams], d["means3D"], d["opacities"],
                       d["scales"], d["rotations"], dc=d["dc"], shs=d["shs"],
                       options=opts).forward()
        torch.cuda.synchronize()
        for a, b in zip(q.images(), bf.images()):
            np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
        np.testing.assert_array_equal(q.T_final.cpu().numpy(), bf.T_final.cpu().numpy())
        np.testing.assert_array_equal(q.last_contributor.cpu().numpy(),
                                      bf.last_contributor.cpu().numpy())
        qb = q.backward(g)
        for v in range(len(cams)):
            raw_close(qb["raw"][v].cpu().numpy(), ref["raw"][v].cpu().numpy(),
                      name=f"queue {rep} raw[{v}]")


def test_tile_shape_batch_matches_single_views(batch_case):
    """On 8 x 32 tiles the batch equals single-view frames on 8 x 32 tiles
    (images, T_final, last contributors, raw sums)."""
    print("\n[multiview] tile_w 8", flush=True)
    scene, cams, d, bf, singles = batch_case
    opts = RasterOptions(schedule=SCHEDULE_FAST, bucket_size=32, tile_w=8)
    b8 = BatchFrame([settings_for(c) for c in cams], d["means3D"], d["opacities"], d["scales"],
                    d["rotations"], dc=d["dc"], shs=d["shs"], options=opts).forward()
    s8 = [Frame(settings_for(c), d["means3D"], d["opacities"], d["scales"], d["rotations"],
                dc=d["dc"], shs=d["shs"], options=opts).forward() for c in cams]
    torch.cuda.synchronize()
    tf, lc = b8.per_view(b8.T_final), b8.per_view(b8.last_contributor)
    g = torch.cat([torch.from_numpy(scenes.target_image(c.width, c.height, seed=v)).cuda()
                   .reshape(-1) for v, c in enumerate(cams)])
    gb = b8.backward(g)
    off = 0
    for v, (img, s) in enumerate(zip(b8.images(), s8)):
        np.testing.assert_array_equal(img.cpu().numpy(), s.out_color.cpu().numpy())
        np.testing.assert_array_equal(tf[v].cpu().numpy(), s.T_final[..., 0].cpu().numpy())
        np.testing.assert_array_equal(lc[v].cpu().numpy(), s.last_contributor[..., 0].cpu().numpy())
        n = img.numel()
        gs = s.backward(g[off:off + n].reshape(img.shape))
        off += n
        raw_close(gb["raw"][v].cpu().numpy(), gs["raw"].cpu().numpy(), name=f"tile8 raw[{v}]")


def test_lists_match_single_views(batch_case):
    scene, cams, d, bf, singles = batch_case
    bv = bf.view_arrays()
    rng = bv["ranges"].cpu().numpy().view(np.uint32).reshape(-1, 2)
    vals = bv["sorted_vals"].cpu().numpy().view(np.uint32)
    P = bf.P
    toff = 0
    for v, s in enumerate(singles):
        sv = s.view_arrays()
        srng = sv["ranges"].cpu().numpy().view(np.uint32).reshape(-1, 2)
        svals = sv["sorted_vals"].cpu().numpy().view(np.uint32)
        T = srng.shape[0]
        for t in range(T):
            a0, a1 = rng[toff + t]
            b0, b1 = srng[t]
            assert a1 - a0 == b1 - b0, (v, t)
            if a1 > a0:
                got = vals[a0:a1].astype(np.int64) - v * P
                np.testing.assert_array_equal(got, svals[b0:b1])
        toff += T
    assert toff == rng.shape[0]


def test_backward_matches_single_views(batch_case):
    print("\n[multiview] backward", flush=True)
    scene, cams, d, bf, singles = batch_case
    targets = [torch.from_numpy(scenes.target_image(c.width, c.height, seed=v)).cuda()
               for v, c in enumerate(cams)]
    grads_flat = torch.cat([(2.0 * (img - t) / img.numel()).reshape(-1)
                            for img, t in zip(bf.images(), targets)])
    gb = bf.backward(grads_flat)
    sums = None
    for v, (s, t) in enumerate(zip(singles, targets)):
        gs = s.backward(2.0 * (s.out_color - t) / s.out_color.numel())
        grad_close(gb["mean2D"][v].cpu().numpy(), gs["mean2D"].cpu().numpy(), 1e-4, f"mean2D[{v}]")
        grad_close(gb["conic"][v].cpu().numpy(), gs["conic"].cpu().numpy(), 1e-4, f"conic[{v}]")
        raw_close(gb["raw"][v].cpu().numpy(), gs["raw"].cpu().numpy(), name=f"raw[{v}]")
        if sums is None:
            sums = {k: gs[k].double().clone() for k in ("means3D", "dc", "shs", "opacities",
                                                         "scales", "rotations", "cov3D")}
        else:
            for k in sums:
                sums[k] += gs[k].double()
    for k, ref in sums.items():
        # the batch adds the views' parameter gradients in float, view order
        grad_close(gb[k].double().cpu().numpy().ravel(), ref.cpu().numpy().ravel(), 1e-4, k)


def test_densify_stats_match_consecutive_views(batch_case):
    scene, cams, d, bf, singles = batch_case
    P = bf.P
    dm = torch.randn(len(singles), P, 4, device="cuda")
    acc_b, accabs_b, mr_b, den_b = (torch.zeros(P, device="cuda") for _ in range(4))
    bf.densify_stats(dm, acc_b, accabs_b, mr_b, den_b)
    acc_s, accabs_s, mr_s, den_s = (torch.zeros(P, device="cuda") for _ in range(4))
    for v, s in enumerate(singles):
        s.densify_stats(dm[v].contiguous(), acc_s, accabs_s, mr_s, den_s)
    for a, b in ((acc_b, acc_s), (accabs_b, accabs_s), (mr_b, mr_s), (den_b, den_s)):
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_autograd_entry_point(batch_case):
    scene, cams, d, bf, singles = batch_case
    params = {k: v.clone().requires_grad_(True) for k, v in d.items()}
    imgs, radii = rasterize_views([settings_for(c) for c in cams], params["means3D"],
                                  params["opacities"], dc=params["dc"], shs=params["shs"],
                                  scales=params["scales"], rotations=params["rotations"],
                                  options=bf.opts)
    assert radii.shape == (len(cams), d["means3D"].shape[0])
    for v, s in enumerate(singles):
        np.testing.assert_array_equal(imgs[v].detach().cpu().numpy(), s.out_color.cpu().numpy())
    loss = sum((img ** 2).mean() for img in imgs)
    loss.backward()
    for k, p in params.items():
        assert p.grad is not None and torch.isfinite(p.grad).all(), k
    assert params["dc"].grad.abs().sum() > 0


def test_autotuner_picks_a_candidate(batch_case, tmp_path):
    from fastergs_amd.autotune import Autotuner
    scene, cams, d, bf, singles = batch_case
    settings = [settings_for(c) for c in cams]
    calls = []

    def step(opts):
        calls.append((opts.bucket_size, opts.split_tiles))
        imgs, _ = rasterize_views(settings, d["means3D"], d["opacities"], dc=d["dc"],
                                  shs=d["shs"], scales=d["scales"], rotations=d["rotations"],
                                  options=opts)
        torch.cuda.synchronize()

    cache = str(tmp_path / "tune.json")
    # bucket size tuned on request (it is semantic); split threshold and the
    # compositing work shape (whole tiles or 16x8 bands) never change outputs
    # and the tile shape on request (semantic as well); the launch form of
    # compositing (one workgroup per item or the work-queue grid) never does
    tuner = Autotuner(buckets=(16, 32), splits=(32, 128), work_splits=(None, 1), reps=1,
                      cache_path=cache, tile_ws=(16, 32))
    best, table = tuner.tune(step, key="k")
    assert len(table) == 32 and table[0][1] <= table[-1][1]
    assert (best.bucket_size, best.split_tiles) in {(16, 32), (16, 128), (32, 32), (32, 128)}
    assert best.tile_w in (16, 32)
    shape = lambda o: None if o.tile_sched is None else o.tile_sched.split_len
    n = len(calls)
    again, _ = Autotuner(cache_path=cache).tune(step, key="k")
    assert len(calls) == n  # served from the cache file
    assert (again.bucket_size, again.split_tiles, shape(again), again.tile_w,
            again.composite_queue) == \
        (best.bucket_size, best.split_tiles, shape(best), best.tile_w, best.composite_queue)
    # without buckets the caller's bucket size is kept
    fixed = Autotuner(splits=(64,), work_splits=(None,), reps=1, queues=(False,),
                      base=RasterOptions(schedule=1, bucket_size=32, tile_w=8))
    assert [(c.bucket_size, c.tile_w) for c in fixed.candidates] == [(32, 8)]
    # the tuned options render exactly what a frame with those options does
    # (the bucket size is part of the reference's semantics -- the T < 1e-4
    # stop is tested once per bucket -- so images of different bucket sizes
    # may differ by up to ~1e-4 of a colour)
    for o in (best, again):
        imgs, _ = rasterize_views(settings, d["means3D"], d["opacities"], dc=d["dc"], shs=d["shs"],
                                  scales=d["scales"], rotations=d["rotations"], options=o)
        same = BatchFrame(settings, d["means3D"], d["opacities"], d["scales"], d["rotations"],
                          dc=d["dc"], shs=d["shs"], options=o).forward()
        for a, b in zip(imgs, same.images()):
            np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
