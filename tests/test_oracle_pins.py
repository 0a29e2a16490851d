"""Independent FP64 pins of oracle functions that the other CPU tests only
checked against the oracle's own intermediate values.

* radius (preprocess.cu:480-486): ceil(min(3.33, sqrt(2 ln(255 op))) *
  sqrt(lambda_max)) from the closed-form 2D covariance of on-axis Gaussians;
* power and the tile rect (preprocess.cu:488-497): the pair set equals the
  tiles whose pixel-centre box meets the FP64 ellipse {q <= 2 mult ln(256
  op)} (equivalently op exp(-q/2) >= 1/256 for mult = 1), by brute force
  over every tile of the image, so a rect that is too small or a wrong
  power both fail;
* the bucket cut-off (render.cu:338-356): a pixel's last contributor is the
  first bucket end at which its FP64 transmittance, blended front to back
  from its own alphas, is below 1e-4 (else its whole list), and T_final is
  that transmittance;
* densification statistics (backward.cu:892-918): the closed form from the
  mean2D gradient, added in place to the caller's tensors.
"""
import math

import numpy as np
import pytest

from fastergs_amd import scenes
from oracle.pyoracle import Oracle


def _axis_scene(n, rng, z=0.0):
    return {
        "means3D": np.tile(np.array([[0.0, 0.0, z]], np.float32), (n, 1)),
        "opacities": rng.uniform(0.002, 1.0, (n, 1)).astype(np.float32),
        "scales": np.exp(rng.uniform(np.log(0.002), np.log(0.3), (n, 3))).astype(np.float32),
        "rotations": np.tile(np.array([[1.0, 0.0, 0.0, 0.0]], np.float32), (n, 1)),
        "dc": np.zeros((n, 1, 3), np.float32),
        "shs": np.zeros((n, 15, 3), np.float32),
    }


def test_radius_closed_form_on_axis():
    rng = np.random.default_rng(21)
    n = 4000
    scene = _axis_scene(n, rng)
    cam = scenes.look_at_camera((0.0, 0.0, -4.0), 640, 480)
    o = Oracle(scene, cam).forward()
    radii = o.get("radii")
    fx = cam.width / (2.0 * cam.tanfovx)
    fy = cam.height / (2.0 * cam.tanfovy)
    s = scene["scales"].astype(np.float64)
    op = scene["opacities"].astype(np.float64).ravel()
    # view depth 4 on the optical axis: J = diag(f / z); view rotation flips x, y
    a = (fx * s[:, 0] / 4.0) ** 2 + 0.3
    c = (fy * s[:, 1] / 4.0) ** 2 + 0.3
    mid = 0.5 * (a + c)
    lam = mid + np.sqrt(np.maximum(0.01, mid * mid - a * c))
    with np.errstate(invalid="ignore", divide="ignore"):
        ext = np.minimum(3.33, np.sqrt(2.0 * np.log(255.0 * op)))
    raw = ext * np.sqrt(lam)
    culled = 255.0 * op < 1.0
    assert culled.sum() >= 5 and (~culled).sum() > 3000
    np.testing.assert_array_equal(radii[culled], 0)
    ok = ~culled & (np.abs(raw - np.round(raw)) > 1e-4)  # away from the ceil boundaries
    assert ok.sum() > 0.99 * (~culled).sum()
    np.testing.assert_array_equal(radii[ok], np.ceil(raw[ok]).astype(np.int32))


def _min_quad_boxes(cx, cy, A, B, C, x0, x1, y0, y1):
    """Vectorised min over closed boxes [x0,x1]x[y0,y1] of A dx^2 + 2B dx dy
    + C dy^2 (d = point - centre), FP64."""
    inside = (x0 <= cx) & (cx <= x1) & (y0 <= cy) & (cy <= y1)
    best = np.full(x0.shape, np.inf)
    for fx in (x0, x1):
        dx = fx - cx
        dy = np.clip(-B * dx / C, y0 - cy, y1 - cy)
        best = np.minimum(best, A * dx * dx + 2 * B * dx * dy + C * dy * dy)
    for fy in (y0, y1):
        dy = fy - cy
        dx = np.clip(-B * dy / A, x0 - cx, x1 - cx)
        best = np.minimum(best, A * dx * dx + 2 * B * dx * dy + C * dy * dy)
    return np.where(inside, 0.0, best)


@pytest.mark.parametrize("mult", [1.0, 1.7, 0.6])
def test_pair_set_from_independent_power(mult):
    scene = scenes.uniform_scene(2500, seed=5, log_mu=-3.6, log_sigma=0.9)
    cam = scenes.look_at_camera((0.2, -0.1, -4.0), 320, 240)
    o = Oracle(scene, cam, mult=mult).forward()
    pj = o.get("proj").reshape(-1, 16)
    keys, vals = o.get("keys"), o.get("vals")
    gx, gy = o.gx, o.gy
    hit = {}
    for t, v in zip((keys >> np.uint64(32)).astype(np.int64), vals):
        hit.setdefault(int(v), set()).add(int(t))
    ty, tx = np.divmod(np.arange(gx * gy), gx)
    x0, x1 = 16.0 * tx, 16.0 * tx + 15.0
    y0, y1 = 16.0 * ty, 16.0 * ty + 15.0
    op = scene["opacities"].astype(np.float64).ravel()
    checked = ambiguous = 0
    for i in np.flatnonzero(pj[:, 0] > 0):
        x, y = float(pj[i, 6]), float(pj[i, 7])
        A, Bc, C = (float(v) for v in pj[i, 9:12])
        thr = 2.0 * mult * math.log(256.0 * op[i])  # not the oracle's power
        q = _min_quad_boxes(x, y, A, Bc, C, x0, x1, y0, y1)
        want = q <= thr
        sure = np.abs(q - thr) > 1e-4 * max(thr, 1e-6)
        got = np.zeros(gx * gy, bool)
        got[list(hit.get(int(i), ()))] = True
        ambiguous += int((~sure).sum())
        checked += int(sure.sum())
        bad = sure & (got != want)
        assert not bad.any(), (i, np.flatnonzero(bad)[:5], thr)
    assert checked > 500000 and ambiguous < 50


def _cutoff_scene():
    """Dense enough that about a third of the pixels reach T < 1e-4."""
    rng = np.random.default_rng(8)
    n = 6000
    scene = scenes.uniform_scene(n, seed=8, radius=0.8, log_mu=-2.3, log_sigma=0.4)
    scene["opacities"] = rng.uniform(0.3, 0.95, (n, 1)).astype(np.float32)
    return scene, scenes.look_at_camera((0.0, 0.0, -2.5), 96, 80)


@pytest.mark.parametrize("bucket", [16, 32])
def test_bucket_cutoff_rule_fp64(bucket):
    scene, cam = _cutoff_scene()
    o = Oracle(scene, cam, bucket=bucket, mode=0).forward()
    pj = o.get("proj").reshape(-1, 16).astype(np.float64)
    op = scene["opacities"].astype(np.float64).ravel()
    vals = o.get("vals").astype(np.int64)
    rng = o.get("ranges").reshape(-1, 2).astype(np.int64)
    contrib = o.get("contrib").reshape(o.H, o.W)
    tfinal = o.get("tfinal").reshape(o.H, o.W)
    n_cut = checked = skipped = 0
    for py in range(o.H):
        for px in range(o.W):
            t = (py // 16) * o.gx + px // 16
            ids = vals[rng[t, 0]:rng[t, 1]]
            if ids.size == 0:
                assert contrib[py, px] == 0 and tfinal[py, px] == 1.0
                checked += 1
                continue
            dx = pj[ids, 6] - px
            dy = pj[ids, 7] - py
            q = pj[ids, 9] * dx * dx + 2 * pj[ids, 10] * dx * dy + pj[ids, 11] * dy * dy
            g = op[ids] * np.exp(-0.5 * q)
            alpha = np.minimum(0.99, g)
            near = (np.abs(g - 1 / 255) < 1e-4 / 255) | (np.abs(g - 0.99) < 1e-6)
            alpha[alpha < 1 / 255] = 0.0
            T = np.cumprod(1.0 - alpha)
            ends = np.minimum(np.arange(bucket, ids.size + bucket, bucket), ids.size)
            Te = T[ends - 1]
            if near.any() or (np.abs(Te / 1e-4 - 1.0) < 1e-3).any():
                skipped += 1
                continue
            below = np.flatnonzero(Te < 1e-4)
            want = int(ends[below[0]]) if below.size else ids.size
            n_cut += bool(below.size)
            assert contrib[py, px] == want, (px, py, contrib[py, px], want)
            np.testing.assert_allclose(tfinal[py, px], T[want - 1], rtol=2e-3, atol=1e-9)
            checked += 1
    assert checked > 0.95 * o.W * o.H and n_cut > 0.2 * checked, (checked, n_cut, skipped)


def test_densify_stats_closed_form():
    scene, cam = scenes.workload("small", n=20000)
    scene["opacities"][:200] = 0.002  # culled (255 op < 1): radii 0, statistics untouched
    o = Oracle(scene, cam, bucket=32, mode=0).forward()
    img = o.image()
    o.backward((img - 0.4).astype(np.float32))
    g = o.grads()["mean2D"].astype(np.float64)
    radii = o.get("radii")
    P = radii.size
    rng = np.random.default_rng(4)
    init = [rng.uniform(0, 1, P).astype(np.float32) for _ in range(4)]
    acc, acc_abs, mr, den = (a.copy() for a in init)
    o.stats(acc, acc_abs, mr, den)
    vis = radii > 0
    assert vis.sum() > 1000 and (~vis).sum() > 10
    e_acc = init[0].astype(np.float64) + np.where(vis, np.hypot(g[:, 0], g[:, 1]), 0.0)
    e_abs = init[1].astype(np.float64) + np.where(vis, np.hypot(g[:, 2], g[:, 3]), 0.0)
    e_mr = np.where(vis, np.maximum(init[2], radii.astype(np.float32)), init[2])
    e_den = init[3] + vis.astype(np.float32)
    np.testing.assert_allclose(acc, e_acc, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(acc_abs, e_abs, rtol=1e-6, atol=1e-12)
    np.testing.assert_array_equal(mr, e_mr)
    np.testing.assert_array_equal(den, e_den)
