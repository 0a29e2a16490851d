"""GPU tests of the drop-in API (GaussianRasterizer / rasterize_gaussians)."""
import numpy as np
import pytest
import torch

from fastergs_amd import scenes
from fastergs_amd.rasterizer import (Frame, GaussianRasterizer, RasterOptions,
                                     get_default_options, set_default_options)
from oracle.pyoracle import Oracle
from tests.helpers import grad_close, settings_for, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[0, 1], ids=["plain", "fast"])
def schedule(request):
    old = get_default_options().schedule
    set_default_options(schedule=request.param)
    yield request.param
    set_default_options(schedule=old)


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
    loss = ((img - tgt) ** 2).mean()
    loss.backward()
    torch.cuda.synchronize()
    mode = 1 if schedule == 0 else 0
    orc = Oracle(scene, cam, bucket=32, mode=mode).forward()
    np.testing.assert_array_equal(img.detach().cpu().numpy().view(np.uint32),
                                  orc.image().view(np.uint32))
    dimg = (2.0 * (orc.image() - scenes.target_image(320, 240)) / img.numel()).astype(np.float32)
    orc.backward(dimg)
    ref = orc.grads()
    for k in ("dc", "shs", "opacities"):
        grad_close(p[k].grad.cpu().numpy(), ref[k], name=k)
    for k in ("means3D", "scales", "rotations"):  # see helpers.check_backward
        a, b = p[k].grad.cpu().numpy().ravel(), ref[k].ravel()
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b), k
    a = np.zeros(P, np.float32); ab = np.zeros(P, np.float32)
    mr = np.zeros(P, np.float32); dn = np.zeros(P, np.float32)
    orc.stats(a, ab, mr, dn)
    np.testing.assert_array_equal(denom.cpu().numpy().ravel(), dn)
    np.testing.assert_array_equal(max_radii.cpu().numpy(), mr)
    for got_s, ref_s in ((accum, a), (accum_abs, ab)):
        g_ = got_s.cpu().numpy().ravel()
        assert np.linalg.norm(g_ - ref_s) <= 1e-4 * np.linalg.norm(ref_s)


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
        opts = RasterOptions(schedule=schedule, speculative=speculative)
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
        for k in ("opacities", "dc", "shs"):
            grad_close(g[k], gref[k], rtol=1e-4, name=k)  # float atomics: run-to-run order
    assert rz._capacity_hints  # the forward remembers the shape's counts


@pytest.mark.parametrize("schedule", [0, 1], ids=["plain", "fast"])
def test_backward_skips_unrequested_gradients(schedule):
    """Skipped outputs (conic, colors, cov3D) are not returned and the others
    equal the full backward's (to float atomics' reassociation)."""
    scene = scenes.uniform_scene(8000, seed=9)
    cam = scenes.look_at_camera((0.0, 0.3, -4.0), 200, 160)
    d = to_dev(scene)
    fr = Frame(settings_for(cam), d["means3D"], d["opacities"], d["scales"], d["rotations"],
               dc=d["dc"], shs=d["shs"], options=RasterOptions(schedule=schedule)).forward()
    gimg = torch.from_numpy(scenes.target_image(200, 160, seed=3)).cuda()
    full = fr.backward(gimg)
    part = fr.backward(gimg, skip=("conic", "colors", "cov3D"))
    assert not {"conic", "colors", "cov3D"} & set(part)
    for k in ("mean2D", "opacities", "dc", "shs"):
        grad_close(part[k].cpu().numpy(), full[k].cpu().numpy(), rtol=1e-5, name=k)
    for k in ("means3D", "scales", "rotations"):  # see helpers.check_backward
        a, b = part[k].cpu().numpy().ravel(), full[k].cpu().numpy().ravel()
        assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b), k
    with pytest.raises(ValueError):
        fr.backward(gimg, skip=("means3D",))
