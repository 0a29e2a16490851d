"""GPU tests of the drop-in API (GaussianRasterizer / rasterize_gaussians)."""
import numpy as np
import pytest
import torch

from fastergs_amd import scenes
from fastergs_amd.rasterizer import (Frame, GaussianRasterizer, RasterOptions,
                                     get_default_options, set_default_options)
from oracle.pyoracle import Oracle
from tests.helpers import rel_l2, grad_close, settings_for, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[0, 1], ids=["plain", "fast"])
def schedule(request):
    old = get_default_options().schedule, get_default_options().bucket_size
    set_default_options(schedule=request.param, bucket_size=32)
    yield request.param
    set_default_options(schedule=old[0], bucket_size=old[1])


def _params(scene):
    d = to_dev(scene)
    return {k: v.clone().requires_grad_(True) for k, v in d.items()}


def test_autograd_training_step(schedule):
    scene = scenes.uniform_scene(20000, seed=5)
    cam = scenes.look_at_camera((0.5, 0.2, -4.0), 320, 240)
    p = _params(scene)
    P = scene["means3D"].shape[0]
    accum = torch.zeros(P, 1, device="cuda")
    accum_abs = torch.zeros(P, 1, device="cuda")
    max_radii = torch.zeros(P, device="cuda")
    denom = torch.zeros(P, 1, device="cuda")
    ras = GaussianRasterizer(settings_for(cam))
    img, radii = ras(means3D=p["means3D"], opacities=p["opacities"], dc=p["dc"], shs=p["shs"],
                     scales=p["scales"], rotations=p["rotations"], xyz_gradient_accum=accum,
                     xyz_gradient_accum_abs=accum_abs, max_radii=max_radii, denom=denom)
    assert img.shape == (3, 240, 320) and radii.dtype == torch.int32
    tgt = torch.from_numpy(scenes.target_image(320, 240)).cuda()
    img.retain_grad()
    loss = ((img - tgt) ** 2).mean()
    loss.backward()
    torch.cuda.synchronize()
    mode = 1 if schedule == 0 else 0
    orc = Oracle(scene, cam, bucket=32, mode=mode).forward()
    np.testing.assert_array_equal(img.detach().cpu().numpy().view(np.uint32),
                                  orc.image().view(np.uint32))
    orc.backward(img.grad.cpu().numpy())  # the dL/dimage autograd passed to the rasterizer
    ref = orc.grads()
    for k in ("dc", "shs", "opacities", "means3D", "scales", "rotations"):
        grad_close(p[k].grad.cpu().numpy(), ref[k], name=k)
        assert rel_l2(p[k].grad.cpu().numpy(), ref[k]) <= 1e-4, k
    a = np.zeros(P, np.float32); ab = np.zeros(P, np.float32)
    mr = np.zeros(P, np.float32); dn = np.zeros(P, np.float32)
    orc.stats(a, ab, mr, dn)
    np.testing.assert_array_equal(denom.cpu().numpy().ravel(), dn)
    np.testing.assert_array_equal(max_radii.cpu().numpy(), mr)
    for got_s, ref_s, name in ((accum, a, "accum"), (accum_abs, ab, "accum_abs")):
        grad_close(got_s.cpu().numpy().ravel(), ref_s, name=name)


def test_colors_precomp_and_culling(schedule):
    scene = scenes.uniform_scene(8000, seed=6)
    cam = scenes.look_at_camera((0.0, 0.0, -4.0), 256, 200)
    rng = np.random.default_rng(0)
    colors = rng.random((8000, 3)).astype(np.float32)
    cull = rng.random(8000) < 0.3
    d = to_dev(scene)
    col = torch.from_numpy(colors).cuda().requires_grad_(True)
    op = d["opacities"].clone().requires_grad_(True)
    ras = GaussianRasterizer(settings_for(cam, bg=(0.2, 0.4, 0.6)))
    img, radii = ras(means3D=d["means3D"], opacities=op, colors_precomp=col, scales=d["scales"],
                     rotations=d["rotations"], culling=torch.from_numpy(cull).cuda())
    (img * img).sum().backward()
    mode = 1 if schedule == 0 else 0
    orc = Oracle(scene, cam, bucket=32, mode=mode, colors=colors, culling=cull,
                 bg=(0.2, 0.4, 0.6)).forward()
    np.testing.assert_array_equal(img.detach().cpu().numpy().view(np.uint32),
                                  orc.image().view(np.uint32))
    assert (radii.cpu().numpy()[cull] == 0).all()
    orc.backward(2.0 * orc.image())
    ref = orc.grads()
    grad_close(col.grad.cpu().numpy(), ref["colors"], name="colors")
    grad_close(op.grad.cpu().numpy(), ref["opacities"], name="opacities")


def test_empty_and_all_culled(schedule):
    cam = scenes.look_at_camera((0.0, 0.0, -4.0), 100, 70)
    scene = scenes.uniform_scene(500, seed=7)
    d = to_dev(scene)
    bg = (0.1, 0.2, 0.3)
    ras = GaussianRasterizer(settings_for(cam, bg=bg))
    # all behind the camera -> culled
    img, radii = ras(means3D=d["means3D"] * 0 + torch.tensor([0.0, 0.0, -10.0], device="cuda"),
                     opacities=d["opacities"], dc=d["dc"], shs=d["shs"], scales=d["scales"],
                     rotations=d["rotations"])
    assert (radii == 0).all()
    exp = torch.tensor(bg, device="cuda").view(3, 1, 1).expand(3, 70, 100)
    assert torch.equal(img, exp)
    # no Gaussians at all
    z = lambda *s: torch.zeros(*s, device="cuda")
    img, radii = ras(means3D=z(0, 3), opacities=z(0, 1), dc=z(0, 1, 3), shs=z(0, 15, 3),
                     scales=z(0, 3), rotations=z(0, 4))
    assert radii.numel() == 0 and torch.equal(img, exp)


def test_argument_errors():
    cam = scenes.look_at_camera((0.0, 0.0, -4.0), 64, 64)
    d = to_dev(scenes.uniform_scene(10, seed=8))
    ras = GaussianRasterizer(settings_for(cam))
    with pytest.raises(ValueError):
        ras(means3D=d["means3D"], opacities=d["opacities"], scales=d["scales"],
            rotations=d["rotations"])  # neither shs nor colours
    with pytest.raises(ValueError):
        ras(means3D=d["means3D"], opacities=d["opacities"], dc=d["dc"], shs=d["shs"],
            cov3D_precomp=torch.zeros(10, 6, device="cuda"))


def test_rasterize_gaussians_with_none_inputs(schedule):
    """Direct calls of rasterize_gaussians may pass None for the unused
    inputs (colors_precomp, cov3Ds_precomp, culling, statistics); backward
    must then return no gradient for them."""
    from fastergs_amd.rasterizer import rasterize_gaussians
    scene = scenes.uniform_scene(5000, seed=8)
    cam = scenes.look_at_camera((0.0, 0.3, -4.0), 160, 120)
    p = _params(scene)
    img, radii = rasterize_gaussians(p["means3D"], p["dc"], p["shs"], None, p["opacities"],
                                     p["scales"], p["rotations"], None, settings_for(cam), None,
                                     None, None, None, None)
    (img ** 2).mean().backward()
    for k in ("means3D", "dc", "shs", "opacities", "scales", "rotations"):
        assert p[k].grad is not None and torch.isfinite(p[k].grad).all(), k


def test_degree0_dc_gradient(schedule):
    """SH degree 0 without rest coefficients: colour = C0 dc + 1/2 (clamped
    at 0), so dL/ddc = C0 dL/dcolour wherever the colour was not clamped."""
    from fastergs_amd.rasterizer import Frame, get_default_options
    from tests.helpers import settings_for as sf
    scene = scenes.uniform_scene(4000, seed=9)
    cam = scenes.look_at_camera((0.0, 0.0, -4.0), 160, 128)
    d = to_dev(scene)
    fr = Frame(sf(cam, sh_degree=0), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=None, options=get_default_options()).forward()
    g = fr.backward(torch.ones_like(fr.out_color))
    C0 = 0.28209479177387814
    col = C0 * d["dc"].reshape(-1, 3) + 0.5
    vis = (fr.radii > 0)[:, None] & (col > 0)
    ddc = g["dc"].reshape(-1, 3)
    dcol = g["colors"].reshape(-1, 3)
    assert vis.sum() > 100
    np.testing.assert_allclose(ddc[vis].cpu().numpy(), (C0 * dcol[vis]).cpu().numpy(),
                               rtol=1e-6, atol=1e-12)
    assert (ddc[~vis] == 0).all()


@pytest.mark.parametrize("schedule", [0, 1], ids=["plain", "fast"])
def test_speculative_capacity_matches_exact(schedule):
    """A forward rendered with a remembered capacity -- larger than needed,
    or too small so that it renders again -- gives the same image, radii and
    gradients as the exact (synchronising) forward."""
    from fastergs_amd import rasterizer as rz
    from fastergs_amd.rasterizer import Frame, RasterOptions
    scene, cam = scenes.uniform_scene(20000, seed=9), scenes.look_at_camera((0.1, 0.2, -4.0), 300, 220)
    d = to_dev(scene)
    dimg = torch.from_numpy(np.random.default_rng(3).normal(size=(3, 220, 300)).astype(np.float32)).cuda()

    def run(speculative, hint=None):
        opts = RasterOptions(schedule=schedule, bucket_size=32, speculative=speculative)
        key = (scene["means3D"].shape[0], 300, 220, schedule, 32, d["means3D"].device)
        rz._capacity_hints.pop(key, None)
        if hint is not None:
            rz._capacity_hints[key] = hint
        fr = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
                   dc=d["dc"], shs=d["shs"], options=opts).forward()
        g = fr.backward(dimg)
        torch.cuda.synchronize()
        return fr, {k: v.cpu().numpy() for k, v in g.items()}

    ref, gref = run(False)
    for hint in [(ref.R * 2 + 5000, ref.NB * 2 + 100), (ref.R // 3, ref.NB // 3), (ref.R, ref.NB)]:
        fr, g = run(True, hint)
        assert (fr.R, fr.NB) == (ref.R, ref.NB)
        assert fr.R_cap >= fr.R and fr.NB_cap >= fr.NB
        np.testing.assert_array_equal(fr.out_color.cpu().numpy(), ref.out_color.cpu().numpy())
        np.testing.assert_array_equal(fr.radii.cpu().numpy(), ref.radii.cpu().numpy())
        np.testing.assert_array_equal(fr.view_arrays()["sorted_vals"].cpu().numpy(),
                                      ref.view_arrays()["sorted_vals"].cpu().numpy())
        for k in ("opacities", "dc", "shs", "means3D", "scales", "rotations"):
            # run-to-run: the float atomics of the linear raw sums reassociate
            grad_close(g[k], gref[k], rtol=1e-5, name=k)
    assert rz._capacity_hints  # the forward remembers the shape's counts


@pytest.mark.parametrize("schedule", [0, 1], ids=["plain", "fast"])
def test_backward_skips_unrequested_gradients(schedule):
    """Skipped outputs (conic, colors, cov3D) are not returned and the others
    equal the full backward's to the reassociation of the float atomics."""
    scene = scenes.uniform_scene(8000, seed=9)
    cam = scenes.look_at_camera((0.0, 0.3, -4.0), 200, 160)
    d = to_dev(scene)
    fr = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=d["shs"], options=RasterOptions(schedule=schedule, bucket_size=32)).forward()
    gimg = torch.from_numpy(scenes.target_image(200, 160, seed=3)).cuda()
    full = fr.backward(gimg)
    part = fr.backward(gimg, skip=("conic", "colors", "cov3D"))
    assert not {"conic", "colors", "cov3D"} & set(part)
    # float atomics of the colour, mean2D and power sums reassociate run to run
    for k in ("mean2D", "opacities", "dc", "shs", "means3D", "scales", "rotations"):
        grad_close(part[k].cpu().numpy(), full[k].cpu().numpy(), rtol=1e-5, name=k)
    with pytest.raises(ValueError):
        fr.backward(gimg, skip=("means3D",))


def test_dropin_defaults_are_reference_semantics():
    """An unchanged caller (no options set) gets FasterGS's bucketed schedule
    with 128-splat buckets (BUCKET_SIZE, rasterizer_imp.h:24): image, radii,
    last contributors and every gradient equal the oracle's bucketed mode at
    B = 128, with the T < 1e-4 cut-off tested only at 128-splat bucket ends."""
    from fastergs_amd._lib import SCHEDULE_FAST
    fresh = RasterOptions()
    assert (fresh.schedule, fresh.bucket_size) == (SCHEDULE_FAST, 128)
    cur = get_default_options()
    assert (cur.schedule, cur.bucket_size) == (SCHEDULE_FAST, 128)
    scene = scenes.cluster_scene(40000, seed=12)
    cam = scenes.look_at_camera((0.2, -0.1, -4.0), 320, 256)
    p = _params(scene)
    ras = GaussianRasterizer(settings_for(cam))
    img, radii = ras(means3D=p["means3D"], opacities=p["opacities"], dc=p["dc"], shs=p["shs"],
                     scales=p["scales"], rotations=p["rotations"])
    tgt = torch.from_numpy(scenes.target_image(320, 256)).cuda()
    img.retain_grad()
    ((img - tgt) ** 2).mean().backward()
    torch.cuda.synchronize()
    orc = Oracle(scene, cam, bucket=128, mode=0).forward()
    assert int(orc.get("max_contrib").max()) > 128  # lists long enough for the bucket size to matter
    np.testing.assert_array_equal(img.detach().cpu().numpy().view(np.uint32),
                                  orc.image().view(np.uint32))
    np.testing.assert_array_equal(radii.cpu().numpy(), orc.get("radii"))
    orc.backward(img.grad.cpu().numpy())
    ref = orc.grads()
    for k in ("dc", "shs", "opacities", "means3D", "scales", "rotations"):
        grad_close(p[k].grad.cpu().numpy(), ref[k], name=k)
        assert rel_l2(p[k].grad.cpu().numpy(), ref[k]) <= 1e-4, k


@pytest.mark.parametrize("schedule", [0, 1], ids=["plain", "fast"])
def test_list_sort_fault_is_reported(schedule):
    """A list sort whose bounded look-back wait gave up marks the frame: the
    image and T_final come out NaN and fgs_read_counts returns an error. The
    condition is forced from the host (the fault word is set between
    preprocess and render); every kernel runs normally."""
    import ctypes
    from fastergs_amd._lib import FasterGSError, FrameView, check, lib
    scene, cam = scenes.uniform_scene(3000, seed=4), scenes.look_at_camera((0.0, 0.0, -4.0), 96, 80)
    d = to_dev(scene)
    fr = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=d["shs"],
               options=RasterOptions(schedule=schedule, bucket_size=32, speculative=False))
    fr.preprocess(sync=True)
    v = FrameView()
    check(lib().fgs_frame_view_get(ctypes.c_void_p(fr.gws.data_ptr()), None, fr.P, fr.W, fr.H, 0,
                                   0, ctypes.byref(fr.copts), ctypes.byref(v)), "view")
    off = v.counts - fr.gws.data_ptr()
    fr.gws[off + 4 * 5] = 1  # the list-sort fault word (COUNT_LIST_FAULT), little endian
    fr.render()
    torch.cuda.synchronize()
    assert torch.isnan(fr.out_color).all() and torch.isnan(fr.T_final).all()
    with pytest.raises(FasterGSError, match="list sort"):
        fr.read_counts(torch.cuda.current_stream())


@pytest.mark.parametrize("schedule", [0, 1], ids=["plain", "fast"])
def test_densify_stats_closed_form(schedule):
    """fgs_densify_stats adds, for every Gaussian with radius > 0, the norm of
    its mean2D gradient (x, y) and of the |.| sums (z, w), one to denom, and
    takes the max with its radius (backward.cu:892-918), in place on the
    caller's values; culled Gaussians keep theirs."""
    scene = scenes.uniform_scene(20000, seed=13)
    scene["opacities"][:300] = 0.002  # culled
    cam = scenes.look_at_camera((0.3, 0.0, -4.0), 320, 240)
    d = to_dev(scene)
    fr = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=d["shs"],
               options=RasterOptions(schedule=schedule, bucket_size=32)).forward()
    g = fr.backward(torch.from_numpy(scenes.target_image(320, 240, seed=5)).cuda())
    P = fr.P
    gen = torch.Generator(device="cuda").manual_seed(3)
    init = [torch.rand(P, device="cuda", generator=gen) for _ in range(4)]
    acc, acc_abs, mr, den = (t.clone() for t in init)
    fr.densify_stats(g["mean2D"], acc, acc_abs, mr, den)
    torch.cuda.synchronize()
    m = g["mean2D"].double().cpu().numpy()
    r = fr.radii.cpu().numpy()
    vis = r > 0
    assert vis.sum() > 1000 and (~vis).sum() >= 300
    i0, i1, i2, i3 = (t.cpu().numpy() for t in init)
    np.testing.assert_allclose(acc.cpu().numpy(), i0 + np.where(vis, np.hypot(m[:, 0], m[:, 1]), 0),
                               rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(acc_abs.cpu().numpy(),
                               i1 + np.where(vis, np.hypot(m[:, 2], m[:, 3]), 0), rtol=1e-6,
                               atol=1e-12)
    np.testing.assert_array_equal(mr.cpu().numpy(), np.where(vis, np.maximum(i2, r.astype(np.float32)), i2))
    np.testing.assert_array_equal(den.cpu().numpy(), i3 + vis.astype(np.float32))
    # the same statistics added by the backward itself (in one pass with the
    # adjoint on FAST frames): bit for bit what fgs_densify_stats adds from
    # that backward's mean2D gradients
    fr2 = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
                dc=d["dc"], shs=d["shs"],
                options=RasterOptions(schedule=schedule, bucket_size=32)).forward()
    st = [t.clone() for t in init]
    g2 = fr2.backward(torch.from_numpy(scenes.target_image(320, 240, seed=5)).cuda(), stats=st)
    sep = [t.clone() for t in init]
    fr2.densify_stats(g2["mean2D"], *sep)
    torch.cuda.synchronize()
    for a_, b_ in zip(st, sep):
        np.testing.assert_array_equal(a_.cpu().numpy(), b_.cpu().numpy())


@pytest.mark.parametrize("schedule", [0, 1], ids=["plain", "fast"])
def test_forward_only_frames(schedule):
    """A frame rendered forward-only (no input needs a gradient, e.g. under
    torch.no_grad) keeps no bucket checkpoints or per-pixel colours, yet its
    image, T_final, contributors and lists equal a full frame's; it cannot be
    differentiated."""
    scene = scenes.cluster_scene(30000, seed=21)
    cam = scenes.look_at_camera((0.1, 0.2, -4.0), 352, 288)
    d = to_dev(scene)
    opts = RasterOptions(schedule=schedule, bucket_size=32)
    full = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
                 dc=d["dc"], shs=d["shs"], options=opts).forward()
    fo = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=d["shs"], options=opts, forward_only=True).forward()
    torch.cuda.synchronize()
    for k in ("out_color", "T_final", "last_contributor", "radii"):
        assert torch.equal(getattr(fo, k), getattr(full, k)), k
    a, b = fo.view_arrays(), full.view_arrays()
    assert torch.equal(a["sorted_vals"], b["sorted_vals"]) and torch.equal(a["ranges"], b["ranges"])
    assert "acc_c" not in a and "pixel_colors" not in a
    assert fo.pws.numel() < full.pws.numel()
    with pytest.raises(ValueError, match="forward-only"):
        fo.backward(torch.ones_like(fo.out_color))
    # the drop-in API: under no_grad the forward is forward-only even for
    # inputs that require gradients (autograd reports them as needing one),
    # with the same image; with grad mode on it is a full frame
    import fastergs_amd.rasterizer as rz
    seen = []

    class Spy(rz.Frame):
        def __init__(self, *a, **k):
            seen.append(bool(k.get("forward_only", False)))
            super().__init__(*a, **k)

    orig = rz.Frame
    rz.Frame = Spy
    try:
        set_default_options(schedule=schedule, bucket_size=32)
        ras = GaussianRasterizer(settings_for(cam))
        p = _params(scene)
        with torch.no_grad():
            img, _ = ras(means3D=p["means3D"], opacities=p["opacities"], dc=p["dc"],
                         shs=p["shs"], scales=p["scales"], rotations=p["rotations"])
        img2, _ = ras(means3D=p["means3D"], opacities=p["opacities"], dc=p["dc"], shs=p["shs"],
                      scales=p["scales"], rotations=p["rotations"])
    finally:
        rz.Frame = orig
    assert seen == [True, False]
    assert torch.equal(img, full.out_color) and torch.equal(img2.detach(), full.out_color)


def test_compositing_zeroes_the_gradient_workspace():
    """fgs_options.zero_grad_ws: the render's compositing zeroes the raw sums
    of a gradient workspace (here filled with NaN bytes first) and the
    backward that takes it with grad_ws_zeroed skips its own zeroing; the
    gradients equal those of a backward that zeroes its own workspace."""
    import ctypes
    from fastergs_amd._lib import lib
    scene, cam = scenes.uniform_scene(20000, seed=21), scenes.look_at_camera((0.2, 0.0, -4.0), 320, 240)
    d = to_dev(scene)

    def frame():
        fr = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
                   dc=d["dc"], shs=d["shs"], options=RasterOptions(bucket_size=32))
        fr.preprocess(sync=True)
        return fr

    dimg = torch.from_numpy(scenes.target_image(cam.width, cam.height, seed=3)).cuda()
    ref = frame().render()
    ref.grad_ws = None  # the backward zeroes a workspace of its own
    g_ref = ref.backward(dimg)
    fr = frame()
    n = lib().fgs_grad_ws_size(fr.P)
    ws = torch.full((n,), 0xFF, dtype=torch.uint8, device="cuda")
    fr.copts.zero_grad_ws = ctypes.c_void_p(ws.data_ptr())
    fr.render()
    fr.copts.zero_grad_ws = None
    torch.cuda.synchronize()
    raw_bytes = fr.P * 64  # conic rows (4 doubles) + the other fields (8 floats)
    assert int(torch.count_nonzero(ws[:raw_bytes])) == 0
    fr.grad_ws = ws
    g = fr.backward(dimg)
    for k in ("opacities", "dc", "shs", "means3D", "scales", "rotations"):
        grad_close(g[k].cpu().numpy(), g_ref[k].cpu().numpy(), name=k)
