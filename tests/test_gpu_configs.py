"""GPU parity at the benchmark workloads' full sizes (one test group per
workload of the benchmark), HIP path against the CPU oracle on the same
seeded inputs:

* small   100k Gaussians, 800x600, bucket 32: both schedules, forward bit for
          bit (radii, pair counts, per-tile lists, ranges, image, T_final, last
          contributors) and the backward (raw sums and every gradient within
          1e-4);
* heavy   1M Gaussians (2% large splats), 1280x840: 21M pairs, above the
          reference's 2^24 pair cap (__init__.py:21-22, :118-119) and above the
          4M keys where the tile sort switches to three-launch passes; FAST
          forward and backward as above, render_simp and render_depth over
          its lists, PLAIN forward, and the drop-in default (FAST, bucket
          128) forward;
* dense   3M Gaussians, 1920x1080 (125M pairs): an orbit frame against the
          oracle, and graph replays of orbit frames equal to eager frames;
* multiview  2M Gaussians, 16 views of uneven sizes in one batch: every view
          equal to its single-view render, two views against the oracle
          (forward and the per-view raw sums of the backward);
* FD      1k Gaussians at 128x128: the HIP adjoint against FP64 central
          finite differences of the FP64 oracle's loss.
"""
import numpy as np
import pytest
import torch

from fastergs_amd import scenes
from fastergs_amd._lib import SCHEDULE_FAST, SCHEDULE_PLAIN
from fastergs_amd.rasterizer import Frame, RasterOptions
from oracle.pyoracle import Oracle
from tests.helpers import check_backward, grad_close, raw_close, settings_for, to_dev

pytestmark = pytest.mark.gpu


def _frame(d, cam, schedule, bucket=32):
    fr = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=d["shs"],
               options=RasterOptions(schedule=schedule, bucket_size=bucket)).forward()
    torch.cuda.synchronize()
    return fr


def _bits(t):
    return t.detach().cpu().numpy().ravel().view(np.uint32)


def assert_forward_equal(fr, orc, lists=True):
    """Forward outputs bit for bit: radii, pair count, lists, ranges, image,
    T_final, last contributors."""
    assert fr.R == orc.get("vals").size
    np.testing.assert_array_equal(fr.radii.cpu().numpy(), orc.get("radii"))
    np.testing.assert_array_equal(_bits(fr.out_color), orc.image().ravel().view(np.uint32))
    np.testing.assert_array_equal(_bits(fr.T_final), orc.get("tfinal").view(np.uint32))
    np.testing.assert_array_equal(fr.last_contributor.cpu().numpy().ravel(), orc.get("contrib"))
    if lists:
        v = fr.view_arrays()
        np.testing.assert_array_equal(v["pair_count"].cpu().numpy().view(np.uint32),
                                      orc.get("pair_count"))
        np.testing.assert_array_equal(v["ranges"].cpu().numpy().view(np.uint32), orc.get("ranges"))
        np.testing.assert_array_equal(v["sorted_vals"].cpu().numpy().view(np.uint32)[:fr.R],
                                      orc.get("vals"))


def _l2_grad(orc, cam):
    img = orc.image()
    return (2.0 * (img - scenes.target_image(cam.width, cam.height)) / img.size).astype(np.float32)


# ------------------------------------------------------------------ small
@pytest.mark.parametrize("schedule", [SCHEDULE_FAST, SCHEDULE_PLAIN], ids=["fast", "plain"])
def test_small_100k_800x600(schedule):
    print(f"\n[config] small, schedule {schedule}", flush=True)
    scene, cam = scenes.workload("small", seed=0)
    fr = _frame(to_dev(scene), cam, schedule)
    orc = Oracle(scene, cam, bucket=32, mode=0 if schedule == SCHEDULE_FAST else 1).forward()
    assert_forward_equal(fr, orc)
    dimg = _l2_grad(orc, cam)
    got = fr.backward(torch.from_numpy(dimg).cuda())
    torch.cuda.synchronize()
    check_backward(got, orc, dimg)


# ------------------------------------------------------------------ heavy
@pytest.fixture(scope="module")
def heavy():
    scene, cam = scenes.workload("heavy", seed=0)
    return scene, cam, to_dev(scene)


def test_heavy_1M_1280x840_fast(heavy):
    print("\n[config] heavy, fast, bucket 32", flush=True)
    scene, cam, d = heavy
    fr = _frame(d, cam, SCHEDULE_FAST)
    assert fr.R > (1 << 24) and fr.R > (1 << 22)  # above the reference's cap and 4M keys
    orc = Oracle(scene, cam, bucket=32, mode=0).forward()
    assert_forward_equal(fr, orc)
    # the list passes over the same lists: render_simp (counts and maxima
    # exact, weight sums within 1e-5) and render_depth (bit for bit); the
    # last tile row (840 = 52.5 tiles) never stops and walks whole lists
    w, c, m = fr.render_simp()
    rw, rc, rm = orc.render_simp()
    np.testing.assert_array_equal(c.cpu().numpy(), rc)
    np.testing.assert_array_equal(m.cpu().numpy(), rm.astype(np.float32))
    grad_close(w.cpu().numpy(), rw, rtol=1e-5, name="accum weights")
    assert rc.sum() > 0 and 0 < rm.sum() <= cam.width * cam.height
    res = fr.render_depth()
    pts, depth, alpha, gidx, disc = orc.render_depth()
    np.testing.assert_array_equal(res["gidx"][0].cpu().numpy(), gidx)
    np.testing.assert_array_equal(res["rendered_depth"][0].cpu().numpy(), depth)
    np.testing.assert_array_equal(res["accum_alpha"][0].cpu().numpy(), alpha)
    del res, w, c, m
    dimg = _l2_grad(orc, cam)
    got = fr.backward(torch.from_numpy(dimg).cuda())
    torch.cuda.synchronize()
    del fr
    check_backward(got, orc, dimg, abs_grad=True)


def test_heavy_1M_1280x840_plain(heavy):
    print("\n[config] heavy, plain", flush=True)
    scene, cam, d = heavy
    fr = _frame(d, cam, SCHEDULE_PLAIN)
    orc = Oracle(scene, cam, bucket=32, mode=1).forward()
    assert_forward_equal(fr, orc)


def test_heavy_1M_1280x840_dropin_default(heavy):
    """The reference's own semantics at the headline size: FAST, 128-splat
    buckets (the drop-in default)."""
    print("\n[config] heavy, fast, bucket 128", flush=True)
    scene, cam, d = heavy
    fr = _frame(d, cam, SCHEDULE_FAST, bucket=128)
    orc = Oracle(scene, cam, bucket=128, mode=0).forward()
    assert_forward_equal(fr, orc, lists=False)


# ------------------------------------------------------------------ dense
def test_dense_3M_1920x1080_frame_and_replays():
    from fastergs_amd.graph import CapturedRenderer
    print("\n[config] dense", flush=True)
    scene, cam = scenes.workload("dense", seed=0)
    d = to_dev(scene)
    cams = scenes.orbit_cameras(120, cam.width, cam.height, radius=3.5)
    fr = _frame(d, cams[30], SCHEDULE_FAST)
    orc = Oracle(scene, cams[30], bucket=32, mode=0).forward()
    assert_forward_equal(fr, orc)
    del orc, fr
    rend = CapturedRenderer(settings_for(cams[0]), d["means3D"], d["opacities"], d["scales"],
                            d["rotations"], dc=d["dc"], shs=d["shs"],
                            options=RasterOptions(schedule=SCHEDULE_FAST, bucket_size=32),
                            headroom=1.25)
    for k in (0, 45, 90, 0):
        c = cams[k]
        cam_t = (torch.from_numpy(c.viewmatrix).cuda(), torch.from_numpy(c.projmatrix).cuda(),
                 torch.from_numpy(c.campos).cuda())
        rend.render(*cam_t)
        if not rend.check(grow=True):  # this view needed more than the capacity: re-captured
            rend.render(*cam_t)
            assert rend.check()
        img = rend.out_color.clone()
        tf = rend.T_final.clone()
        torch.cuda.synchronize()
        ref = _frame(d, c, SCHEDULE_FAST)  # eager work between replays
        assert torch.equal(img, ref.out_color) and torch.equal(tf, ref.T_final), k
        del ref


# -------------------------------------------------------------- multiview
def test_multiview_2M_16_views():
    from fastergs_amd.multiview import BatchFrame
    print("\n[config] multiview", flush=True)
    scene = scenes.cluster_scene(2_000_000, seed=0)
    cams = scenes.sphere_cameras(16, seed=0)
    d = to_dev(scene)
    opts = RasterOptions(schedule=SCHEDULE_FAST, bucket_size=32)
    bf = BatchFrame([settings_for(c) for c in cams], d["means3D"], d["opacities"], d["scales"],
                    d["rotations"], dc=d["dc"], shs=d["shs"], options=opts).forward()
    torch.cuda.synchronize()
    imgs = bf.images()
    tfs = bf.per_view(bf.T_final)
    lcs = bf.per_view(bf.last_contributor)
    for v, c in enumerate(cams):
        s = _frame(d, c, SCHEDULE_FAST)
        assert torch.equal(imgs[v], s.out_color), v
        assert torch.equal(tfs[v], s.T_final.reshape(c.height, c.width)), v
        assert torch.equal(lcs[v], s.last_contributor.reshape(c.height, c.width)), v
        del s
    # two views against the oracle: forward, and the per-view raw sums of the
    # batch backward for an L2 loss of every view
    targets = [scenes.target_image(c.width, c.height, seed=v) for v, c in enumerate(cams)]
    gflat = torch.cat([(2.0 * (img - torch.from_numpy(t).cuda()) / img.numel()).reshape(-1)
                       for img, t in zip(imgs, targets)])
    g = bf.backward(gflat, skip=("conic", "colors", "cov3D"))
    torch.cuda.synchronize()
    order = sorted(range(len(cams)), key=lambda v: cams[v].width * cams[v].height)
    for v in (order[0], order[-1]):
        c = cams[v]
        print(f"[config] multiview oracle view {v} ({c.width}x{c.height})", flush=True)
        orc = Oracle(scene, c, bucket=32, mode=0).forward()
        np.testing.assert_array_equal(imgs[v].cpu().numpy().view(np.uint32),
                                      orc.image().view(np.uint32))
        np.testing.assert_array_equal(lcs[v].cpu().numpy().ravel(), orc.get("contrib"))
        np.testing.assert_array_equal(bf.radii[v].cpu().numpy(), orc.get("radii"))
        dimg = gflat[3 * bf.pix_off[v]:3 * bf.pix_off[v + 1]].cpu().numpy()
        orc.backward(dimg)
        raw_close(g["raw"][v].cpu().numpy(), orc.get("raw").reshape(-1, 11), name=f"raw[{v}]")
        del orc


# --------------------------------------------------------------------- FD
def _loss64(scene, cam, w, mode):
    o = Oracle(scene, cam, bucket=32, mode=mode, dtype=np.float64).forward()
    return float((o.image() * w).sum())


@pytest.mark.parametrize("schedule", [SCHEDULE_FAST, SCHEDULE_PLAIN], ids=["fast", "plain"])
def test_fd_1k_128x128_hip_adjoint(schedule):
    """The HIP adjoint (float32, on the GPU) of L = sum(image * w) against
    (a) the FP64 oracle's adjoint at the same parameter values, every entry,
    and (b) FP64 central finite differences of the FP64 oracle's loss on
    sampled entries (entries whose two step sizes disagree cross a threshold
    -- alpha 1/255, the T cut-off, a tile rect -- and are skipped). Both within
    1e-4 relative to |reference| + rms of the gradient tensor: float32
    rounding scales with a Gaussian's whole gradient, so a small component of
    a large gradient vector is not held to its own magnitude."""
    print(f"\n[config] FD, schedule {schedule}", flush=True)
    mode = 0 if schedule == SCHEDULE_FAST else 1
    scene, cam = scenes.fd_scene(1000, seed=0)
    w = np.random.default_rng(1).standard_normal((3, cam.height, cam.width))
    fr = _frame(to_dev(scene), cam, schedule)
    g = fr.backward(torch.from_numpy(w.astype(np.float32)).cuda())
    torch.cuda.synchronize()
    names = ("means3D", "scales", "rotations", "opacities", "dc", "shs")
    grads = {k: g[k].cpu().numpy().reshape(scene[k].shape) for k in names}
    s64 = {k: v.astype(np.float64) for k, v in scene.items()}
    o64 = Oracle(s64, cam, bucket=32, mode=mode, dtype=np.float64).forward()
    o64.backward(w)
    ref64 = o64.grads()
    for k in names:
        grad_close(grads[k], ref64[k], rtol=1e-4, name=f"{k} vs the FP64 adjoint")
    radii = fr.radii.cpu().numpy()
    vis = np.flatnonzero(radii > 0)
    rng = np.random.default_rng(7)
    ok = skipped = 0
    worst = 0.0
    for name in names:
        rms = float(np.sqrt(np.mean(ref64[name].astype(np.float64) ** 2)))
        for _ in range(12):
            i = int(rng.choice(vis))
            arr = scene[name]
            idx = (i, *np.unravel_index(rng.integers(arr[i].size), arr[i].shape))
            x0 = float(arr[idx])
            fds = []
            for h in (1e-6, 5e-7):
                sp = {k: v.astype(np.float64) for k, v in scene.items()}
                sp[name][idx] = x0 + h
                lp = _loss64(sp, cam, w, mode)
                sp[name][idx] = x0 - h
                lm = _loss64(sp, cam, w, mode)
                fds.append((lp - lm) / (2 * h))
            an = float(grads[name][idx])
            if abs(fds[0] - fds[1]) > 1e-3 * max(abs(fds[0]), abs(an), 1e-9):
                skipped += 1
                continue
            bound = 1e-4 * (abs(fds[0]) + rms)
            worst = max(worst, abs(fds[0] - an) / (abs(fds[0]) + rms))
            assert abs(fds[0] - an) <= bound, (name, idx, fds, an, rms)
            ok += 1
    print(f"[config] FD: {ok} entries within 1e-4 (worst {worst:.2e}), {skipped} skipped",
          flush=True)
    assert ok >= 60 and skipped <= 8
