"""GPU parity of the PLAIN schedule against the CPU oracle (sequential mode).

Bit-exact: radii, pair counts, per-Gaussian records, sorted keys / values,
tile ranges, images, T_final, last contributors. Gradients: within 1e-4
(relative, see helpers.grad_close) because float atomics reassociate sums.
"""
import ctypes

import numpy as np
import pytest
import torch

from fastergs_amd import scenes
from fastergs_amd._lib import SCHEDULE_PLAIN, check, lib
from fastergs_amd.rasterizer import Frame, RasterOptions
from oracle.pyoracle import Oracle
from tests.helpers import check_backward, grad_close, settings_for, to_dev

pytestmark = pytest.mark.gpu

CASES = {
    "uniform_30k_400x300": lambda: (scenes.uniform_scene(30000, seed=1),
                                    scenes.look_at_camera((0.0, 0.0, -4.0), 400, 300)),
    "clusters_40k_640x420": lambda: (scenes.cluster_scene(40000, seed=2),
                                     scenes.look_at_camera((0.3, -0.2, -4.5), 640, 420)),
}


def run_pair(scene, cam, schedule=SCHEDULE_PLAIN, bucket=32, mode=1, tile_w=16, **kw):
    d = to_dev(scene)
    fr = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=d["shs"], options=RasterOptions(schedule=schedule,
                                                               bucket_size=bucket,
                                                               tile_w=tile_w)).forward()
    torch.cuda.synchronize()
    orc = Oracle(scene, cam, bucket=bucket, mode=mode, tile_w=tile_w, **kw).forward()
    return fr, orc


@pytest.mark.parametrize("tile_w", [8, 32])
def test_tile_shapes(tile_w):
    """On 8 x 32 and 32 x 8 tiles the PLAIN path matches the oracle on the
    same tiles bit for bit (sorted pairs, ranges, image, T_final, last
    contributors). Against 16 x 16 tiles: a Gaussian left out of a tile's
    list has alpha < 1/255 at every pixel of the tile, so a pixel that never
    reaches T < 1e-4 blends the same splats in the same order on every tile
    shape (identical colour and T_final); a pixel that does stops at a
    bucket end of its own tile's list, so the two differ by less than the
    transmittance left at the earlier stop (< 1e-4 of a colour)."""
    print(f"\n[plain] tile_w {tile_w}", flush=True)
    scene, cam = CASES["clusters_40k_640x420"]()
    fr, orc = run_pair(scene, cam, tile_w=tile_w)
    v = fr.view_arrays()
    np.testing.assert_array_equal(v["sorted_keys"].cpu().numpy().view(np.uint64), orc.get("keys"))
    np.testing.assert_array_equal(v["sorted_vals"].cpu().numpy().view(np.uint32), orc.get("vals"))
    np.testing.assert_array_equal(v["ranges"].cpu().numpy().view(np.uint32), orc.get("ranges"))
    np.testing.assert_array_equal(fr.out_color.cpu().numpy().view(np.uint32),
                                  orc.image().view(np.uint32))
    np.testing.assert_array_equal(fr.last_contributor.cpu().numpy().ravel(), orc.get("contrib"))
    ref, _ = run_pair(scene, cam, tile_w=16)
    img, rimg = fr.out_color.cpu().numpy(), ref.out_color.cpu().numpy()
    tf, rtf = fr.T_final.cpu().numpy()[..., 0], ref.T_final.cpu().numpy()[..., 0]
    open_px = (rtf >= 1e-4) & (tf >= 1e-4)
    assert open_px.mean() > 0.2
    np.testing.assert_array_equal(img[:, open_px], rimg[:, open_px])
    np.testing.assert_array_equal(tf[open_px], rtf[open_px])
    cut = rimg.max() * 1e-4
    assert np.abs(img - rimg).max() <= cut, (np.abs(img - rimg).max(), cut)


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    print(f"\n[case] {request.param}", flush=True)
    scene, cam = CASES[request.param]()
    fr, orc = run_pair(scene, cam)
    return scene, cam, fr, orc


def test_counts_and_radii(case):
    scene, cam, fr, orc = case
    assert fr.R == orc.get("vals").size
    np.testing.assert_array_equal(fr.radii.cpu().numpy(), orc.get("radii"))
    v = fr.view_arrays()
    np.testing.assert_array_equal(v["pair_count"].cpu().numpy().view(np.uint32),
                                  orc.get("pair_count"))


def test_records_bit_exact(case):
    scene, cam, fr, orc = case
    v = fr.view_arrays()
    pj = orc.get("proj").reshape(-1, 16)
    valid = pj[:, 0] > 0
    geom = v["geom"].cpu().numpy().reshape(-1, 8)[valid]
    spl = v["splat"].cpu().numpy().reshape(-1, 12)[valid]
    p = pj[valid]
    np.testing.assert_array_equal(geom[:, 0].view(np.uint32), p[:, 6].view(np.uint32))  # x
    np.testing.assert_array_equal(geom[:, 1].view(np.uint32), p[:, 7].view(np.uint32))  # y
    np.testing.assert_array_equal(geom[:, 2:5].view(np.uint32), p[:, 9:12].view(np.uint32))
    np.testing.assert_array_equal(geom[:, 5].view(np.uint32), p[:, 12].view(np.uint32))
    np.testing.assert_array_equal(spl[:, 6].view(np.uint32), p[:, 8].view(np.uint32))  # depth
    np.testing.assert_array_equal(spl[:, 8:11].view(np.uint32), p[:, 13:16].view(np.uint32))
    rmin = geom[:, 6].view(np.uint32)
    rmax = geom[:, 7].view(np.uint32)
    np.testing.assert_array_equal(rmin & 0xFFFF, p[:, 1].astype(np.uint32))
    np.testing.assert_array_equal(rmin >> 16, p[:, 2].astype(np.uint32))
    np.testing.assert_array_equal(rmax & 0xFFFF, p[:, 3].astype(np.uint32))
    np.testing.assert_array_equal(rmax >> 16, p[:, 4].astype(np.uint32))
    np.testing.assert_array_equal(spl[:, 11].view(np.uint32), p[:, 5].astype(np.uint32))


def test_sorted_pairs_and_ranges(case):
    scene, cam, fr, orc = case
    v = fr.view_arrays()
    np.testing.assert_array_equal(v["sorted_keys"].cpu().numpy().view(np.uint64), orc.get("keys"))
    np.testing.assert_array_equal(v["sorted_vals"].cpu().numpy().view(np.uint32), orc.get("vals"))
    np.testing.assert_array_equal(v["ranges"].cpu().numpy().view(np.uint32), orc.get("ranges"))


def test_image_bit_exact(case):
    scene, cam, fr, orc = case
    np.testing.assert_array_equal(fr.out_color.cpu().numpy().view(np.uint32),
                                  orc.image().view(np.uint32))
    np.testing.assert_array_equal(fr.T_final.cpu().numpy().ravel().view(np.uint32),
                                  orc.get("tfinal").view(np.uint32))
    np.testing.assert_array_equal(fr.last_contributor.cpu().numpy().ravel(), orc.get("contrib"))
    v = fr.view_arrays()
    np.testing.assert_array_equal(v["max_contrib"].cpu().numpy().view(np.uint32),
                                  orc.get("max_contrib"))


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


def _sort_case(n, bits, seed, with_vals=True, ws=None, distinct=None):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 2 ** bits, n, dtype=np.uint64, endpoint=False)
    if distinct is not None:  # few distinct digits: long runs per digit
        keys = rng.choice(keys[:distinct], n)
    keys[rng.random(n) < 0.3] = keys[0]  # many equal keys: stability matters
    vals = np.arange(n, dtype=np.uint32)
    k = torch.from_numpy(keys.view(np.int64)).cuda()
    ka = torch.empty_like(k)
    v = torch.from_numpy(vals.view(np.int32)).cuda() if with_vals else None
    va = torch.empty_like(v) if with_vals else None
    L = lib()
    if ws is None:
        ws = torch.empty(L.fgs_sort_ws_size(n), dtype=torch.uint8, device="cuda")
    p = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(L.fgs_sort_pairs(p(k), p(v), p(ka), p(va), n, 0, bits, p(ws), ws.numel(), stream),
          "sort")
    check(L.fgs_sort_check(p(ws), n, stream), "sort check")  # no look-back wait gave up
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k.cpu().numpy().view(np.uint64), keys[order])
    if with_vals:
        np.testing.assert_array_equal(v.cpu().numpy().view(np.uint32), vals[order])


@pytest.mark.parametrize("n,bits", [(1, 45), (4095, 45), (4097, 13), (100003, 45), (1 << 20, 48)])
def test_radix_sort_stable(n, bits):
    _sort_case(n, bits, seed=n)


def test_radix_sort_keys_only():
    _sort_case(70001, 40, seed=3, with_vals=False)


def test_radix_sort_all_passes_and_few_digits():
    _sort_case(3_000_001, 64, seed=11)                 # eight passes
    _sort_case(4_500_000, 45, seed=13)                 # above 4M keys: three-launch passes
    _sort_case(2_000_000, 48, seed=12, distinct=5)      # a handful of digits per pass


def test_radix_sort_workspace_reuse():
    # the look-back status rows are cleared by the sort itself: a second call
    # on the same workspace (smaller input) must not see the first one's state
    ws = torch.empty(lib().fgs_sort_ws_size(300_000), dtype=torch.uint8, device="cuda")
    _sort_case(300_000, 45, seed=21, ws=ws)
    _sort_case(180_000, 45, seed=22, ws=ws)
    _sort_case(300_000, 13, seed=23, ws=ws)


@pytest.mark.parametrize("n", [0, 1, 2047, 2048, 2049, 5_000_000])
def test_scan(n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 100, n, dtype=np.uint32)
    xi = torch.from_numpy(x.view(np.int32)).cuda()
    out = torch.empty_like(xi)
    tot = torch.zeros(1, dtype=torch.int32, device="cuda")
    L = lib()
    ws = torch.empty(L.fgs_scan_ws_size(n), dtype=torch.uint8, device="cuda")
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    check(L.fgs_scan_u32(p(xi), p(out), n, p(tot), p(ws), ws.numel(),
                         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "scan")
    ex = np.concatenate([[0], np.cumsum(x.astype(np.uint64))[:-1]]).astype(np.uint32) if n else x
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), ex)
    assert int(tot.item()) == int(x.astype(np.uint64).sum())
