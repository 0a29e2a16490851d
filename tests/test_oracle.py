"""CPU tests that pin the oracle (no GPU).

The reference repository ships no tests or golden vectors, so each oracle
function is pinned by something it does not share code with: libm for
exp2/log2, closed forms for projection and covariances, FP64 brute force for
the tile test, definitions for sort / ranges / buckets, an analytic image
for one Gaussian, FP64 central finite differences for the adjoint and an
independent PyTorch FP64 autograd restatement (tests/test_oracle_autograd.py).
"""
import math

import numpy as np
import pytest

from fastergs_amd import scenes
from oracle import pyoracle
from oracle.pyoracle import Oracle


def ulps(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.spacing(np.abs(b).astype(np.float32)).astype(np.float64)


def test_exp2_accuracy():
    xs = np.concatenate([np.linspace(-125.0, 10.0, 20001), np.linspace(-9.0, 0.0, 20001),
                         [-0.5, 0.5, 0.0, -1.0, -7.99435]]).astype(np.float32)
    got = np.array([pyoracle.exp2f(float(x)) for x in xs], np.float32)
    ref = np.exp2(xs.astype(np.float64))
    assert ulps(got, ref).max() <= 1.0
    assert pyoracle.exp2f(-200.0) == 0.0
    assert pyoracle.exp2f(0.0) == 1.0


def test_log2_accuracy():
    xs = np.concatenate([np.geomspace(1e-30, 1e30, 20001), np.linspace(1 / 255, 1.0, 20001),
                         np.linspace(1.0, 255.0, 2001)]).astype(np.float32)
    got = np.array([pyoracle.log2f(float(x)) for x in xs], np.float32)
    ref = np.log2(xs.astype(np.float64))
    nz = np.abs(ref) > 1e-3
    assert ulps(got[nz], ref[nz]).max() <= 3.0
    assert np.abs(got[~nz] - ref[~nz]).max() < 1e-9
    assert pyoracle.log2f(1.0) == 0.0 and pyoracle.log2f(8.0) == 3.0


def _one(means, scales, quat=(1.0, 0.0, 0.0, 0.0), op=0.5, W=256, H=192, dc=(0.3, -0.2, 0.1),
         mode=0, **kw):
    n = len(means)
    scene = {
        "means3D": np.asarray(means, np.float32).reshape(n, 3),
        "opacities": np.full((n, 1), op, np.float32),
        "scales": np.asarray(scales, np.float32).reshape(n, 3),
        "rotations": np.tile(np.asarray(quat, np.float32), (n, 1)),
        "dc": np.tile(np.asarray(dc, np.float32), (n, 1)).reshape(n, 1, 3),
        "shs": np.zeros((n, 15, 3), np.float32),
    }
    cam = scenes.look_at_camera((0.0, 0.0, -4.0), W, H)
    return scene, cam, Oracle(scene, cam, mode=mode, **kw).forward()


def test_closed_form_projection_and_cov2d():
    sx, sy = 0.02, 0.035
    scene, cam, o = _one([[0.0, 0.0, 0.0], [0.3, -0.2, 0.5]], [[sx, sy, 0.01]] * 2)
    pj = o.get("proj").reshape(-1, 16)
    fx = cam.width / (2 * cam.tanfovx)
    fy = cam.height / (2 * cam.tanfovy)
    # on-axis Gaussian at view depth 4: cov2D = diag(f^2 s^2 / z^2) + 0.3
    a = (fx * sx / 4.0) ** 2 + 0.3
    c = (fy * sy / 4.0) ** 2 + 0.3
    np.testing.assert_allclose(pj[0, 9:12], [1 / a, 0.0, 1 / c], rtol=2e-6, atol=1e-9)
    np.testing.assert_allclose(pj[0, 6:9], [(cam.width - 1) / 2, (cam.height - 1) / 2, 4.0],
                               rtol=1e-6)
    # off-axis point: pixel = f X/Z + (W-1)/2 with view X = -world x, Y = -world y
    X, Y, Z = -0.3, 0.2, 4.5
    np.testing.assert_allclose(pj[1, 6:8], [fx * X / Z + (cam.width - 1) / 2,
                                            fy * Y / Z + (cam.height - 1) / 2], rtol=1e-5)
    # SH degree 0 colour (ops.h:17, preprocess.cu:132,158): C0*dc + 0.5
    np.testing.assert_allclose(pj[0, 13:16], 0.28209479177387814 * np.array([0.3, -0.2, 0.1]) + 0.5,
                               rtol=1e-6)


def test_rotation_invariance_of_isotropic_cov():
    rng = np.random.default_rng(0)
    q = rng.standard_normal((8, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    outs = []
    for qi in q:
        _, _, o = _one([[0.1, 0.1, 0.0]], [[0.03, 0.03, 0.03]], quat=qi)
        outs.append(o.get("proj").reshape(-1, 16)[0, 9:12])
    np.testing.assert_allclose(np.array(outs), np.tile(outs[0], (8, 1)), rtol=2e-5, atol=5e-7)


def _min_quad_on_box(cx, cy, A, B, C, x0, x1, y0, y1):
    """min over the closed box of q(d) = A dx^2 + 2B dx dy + C dy^2, d = p - centre (FP64)."""
    if x0 <= cx <= x1 and y0 <= cy <= y1:
        return 0.0
    best = math.inf
    for fixed_x in (x0, x1):  # vertical edges: minimise over y in [y0, y1]
        dx = fixed_x - cx
        dy = np.clip(-B * dx / C, y0 - cy, y1 - cy)
        best = min(best, A * dx * dx + 2 * B * dx * dy + C * dy * dy)
    for fixed_y in (y0, y1):
        dy = fixed_y - cy
        dx = np.clip(-B * dy / A, x0 - cx, x1 - cx)
        best = min(best, A * dx * dx + 2 * B * dx * dy + C * dy * dy)
    return best


@pytest.mark.parametrize("tile_w", [16, 8, 32])
def test_tile_set_brute_force(tile_w):
    """Every Gaussian's tiles are exactly the tiles (of each tile shape) whose
    pixel-centre box meets the ellipse {q <= 2 power}, by a minimisation of
    the quadratic form over each box."""
    tw, th = tile_w, 256 // tile_w
    scene = scenes.uniform_scene(3000, seed=3, log_mu=-3.5, log_sigma=0.8)
    cam = scenes.look_at_camera((0.2, 0.1, -4.0), 320, 240)
    o = Oracle(scene, cam, tile_w=tile_w).forward()
    pj = o.get("proj").reshape(-1, 16)
    keys, vals = o.get("keys"), o.get("vals")
    tiles = (keys >> np.uint64(32)).astype(np.int64)
    gx = o.gx
    hit = {}
    for t, v in zip(tiles, vals):
        hit.setdefault(int(v), set()).add(int(t))
    checked = ambiguous = 0
    for i in np.flatnonzero(pj[:, 0] > 0)[:1500]:
        x, y = float(pj[i, 6]), float(pj[i, 7])
        A, B, C = (float(v) for v in pj[i, 9:12])
        power = float(pj[i, 12])
        r0, r1, r2, r3 = (int(v) for v in pj[i, 1:5])
        got = hit.get(int(i), set())
        assert all(r0 <= t % gx < r2 and r1 <= t // gx < r3 for t in got)  # subset of rect
        cx_t, cy_t = int(x // tw), int(y // th)
        if 0 <= x < o.W and 0 <= y < o.H and r0 <= cx_t < r2 and r1 <= cy_t < r3:
            assert cy_t * gx + cx_t in got  # centre tile always present
        for ty in range(r1, r3):
            for tx in range(r0, r2):
                q = _min_quad_on_box(x, y, A, B, C, tw * tx, tw * tx + tw - 1, th * ty,
                                     th * ty + th - 1)
                thr = 2 * power
                if abs(q - thr) <= 1e-4 * thr:
                    ambiguous += 1
                    continue
                checked += 1
                assert ((ty * gx + tx) in got) == (q <= thr), (i, tx, ty, q, thr)
    assert checked > 4000 and ambiguous < checked * 1e-3


def test_factored_tile_test_equals_tile_hit():
    """The per-column / per-row factoring of the exact tile test (used by the
    GPU's walk over large rects) gives tile_hit's answer on every tile, for
    ellipses from round to needle-thin at every orientation, centred inside
    and outside the grid, small and large."""
    import ctypes
    L = pyoracle.lib()
    rng = np.random.default_rng(17)
    hits = ctypes.c_int(0)
    total_hits = 0
    for _ in range(3000):
        s1 = float(np.exp(rng.uniform(np.log(0.3), np.log(600.0))))
        s2 = float(s1 * np.exp(rng.uniform(np.log(0.001), 0.0)))
        th = rng.uniform(0, np.pi)
        R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        con = np.linalg.inv(R @ np.diag([s1 * s1, s2 * s2]) @ R.T + 0.3 * np.eye(2))
        power = float(np.log(256.0 * rng.uniform(0.005, 1.0)))
        x, y = (float(v) for v in rng.uniform(-300, 1600, 2))
        for tw in (16, 8, 32):
            bad = L.orc_tile_terms_check(x, y, con[0, 0], con[0, 1], con[1, 1], power,
                                         1600 // tw, 1120 * tw // 256, ctypes.byref(hits), tw)
            assert bad == 0, (x, y, con, power, tw)
            total_hits += hits.value
    assert total_hits > 100000


@pytest.mark.parametrize("bucket", [16, 32, 128])
def test_sort_ranges_buckets_invariants(bucket):
    scene, cam = scenes.workload("small", n=20000)
    o = Oracle(scene, cam, bucket=bucket).forward()
    keys, vals = o.get("keys"), o.get("vals")
    R = keys.size
    order = np.lexsort((vals, keys))
    np.testing.assert_array_equal(order, np.arange(R))  # sorted by (key, index)
    assert np.array_equal(np.bincount(vals, minlength=o.P), o.get("pair_count"))
    tiles = (keys >> np.uint64(32)).astype(np.int64)
    rng = o.get("ranges").reshape(-1, 2).astype(np.int64)
    T = rng.shape[0]
    cnt = np.bincount(tiles, minlength=T)
    np.testing.assert_array_equal(rng[:, 1] - rng[:, 0], cnt)
    nz = cnt > 0
    np.testing.assert_array_equal(rng[nz, 0], np.searchsorted(tiles, np.arange(T))[nz])
    nb = (cnt + bucket - 1) // bucket
    np.testing.assert_array_equal(o.get("bucket_off"), np.cumsum(nb))
    br = o.get("bucket_ranges").reshape(-1, 2).astype(np.int64)
    bt = o.get("bucket_tile").astype(np.int64)
    lens = br[:, 1] - br[:, 0]
    assert lens.min() >= 1 and lens.max() <= bucket
    for t in np.flatnonzero(nz)[:200]:
        sel = bt == t
        assert lens[sel][:-1].tolist() == [bucket] * (sel.sum() - 1)
        assert br[sel, 0].min() == rng[t, 0] and br[sel, 1].max() == rng[t, 1]


def test_single_gaussian_analytic_image():
    op = 0.7
    scene, cam, o = _one([[0.05, -0.03, 0.2]], [[0.06, 0.03, 0.02]],
                         quat=(0.9, 0.2, -0.3, 0.1), op=op, W=96, H=80, bg=(0.1, 0.2, 0.3))
    pj = o.get("proj").reshape(-1, 16)[0]
    x, y = pj[6], pj[7]
    A, B, C = pj[9:12].astype(np.float64)
    rgb = pj[13:16].astype(np.float64)
    yy, xx = np.mgrid[0:80, 0:96].astype(np.float64)
    dx, dy = x - xx, y - yy
    alpha = np.minimum(0.99, op * np.exp(-0.5 * (A * dx * dx + 2 * B * dx * dy + C * dy * dy)))
    alpha[alpha < 1 / 255] = 0
    exp = rgb[:, None, None] * alpha + (1 - alpha) * np.array([0.1, 0.2, 0.3])[:, None, None]
    np.testing.assert_allclose(o.image(), exp, atol=2e-6)


def test_compositing_invariants_and_modes():
    scene, cam = scenes.workload("small", n=30000)
    ob = Oracle(scene, cam, bucket=32, mode=0).forward()
    os_ = Oracle(scene, cam, bucket=32, mode=1).forward()
    T = ob.get("tfinal")
    assert (T >= 0).all() and (T <= 1).all()
    diff = np.abs(ob.image() - os_.image())
    flips = (ob.get("contrib") != os_.get("contrib")).sum()
    assert flips <= 2
    same = (ob.get("contrib") == os_.get("contrib")).reshape(ob.H, ob.W)
    assert diff[:, same].max() <= 1e-5  # FP32 reassociation between bucket merge and sequential
    # contributors stop at a bucket boundary or at the end of the tile list
    c = ob.get("contrib").reshape(ob.H, ob.W)
    rng = ob.get("ranges").reshape(-1, 2)
    lens = (rng[:, 1] - rng[:, 0]).reshape(ob.gy, ob.gx)
    L = np.repeat(np.repeat(lens, 16, 0), 16, 1)[:ob.H, :ob.W]
    assert ((c % 32 == 0) | (c == L)).all()


def test_merge_is_associative():
    rng = np.random.default_rng(1)
    def op(a, b):
        return (a[0] * b[0], a[1] + a[0] * b[1])
    for _ in range(100):
        u, v, w = [(rng.random(), rng.random(3)) for _ in range(3)]
        l = op(op(u, v), w)
        r = op(u, op(v, w))
        np.testing.assert_allclose(l[0], r[0], rtol=1e-12)
        np.testing.assert_allclose(l[1], r[1], rtol=1e-12)


def _fd_setup(n=1000, seed=0):
    scene, cam = scenes.fd_scene(n, seed)
    w = np.random.default_rng(seed + 1).standard_normal((3, cam.height, cam.width))
    return scene, cam, w


def _loss64(scene, cam, w, mode):
    o = Oracle(scene, cam, bucket=32, mode=mode, dtype=np.float64).forward()
    return float((o.image() * w).sum())


@pytest.mark.parametrize("mode", [0, 1])
def test_fp64_finite_differences(mode):
    scene, cam, w = _fd_setup()
    o = Oracle(scene, cam, bucket=32, mode=mode, dtype=np.float64).forward()
    o.backward(w)
    g = o.grads()
    radii = o.get("radii")
    rng = np.random.default_rng(5)
    vis = np.flatnonzero(radii > 0)
    names = {"means3D": "means3D", "scales": "scales", "rotations": "rotations",
             "opacities": "opacities", "dc": "dc", "shs": "shs"}
    ok = skipped = 0
    for name in names:
        for _ in range(12):
            i = int(rng.choice(vis))
            arr = scene[name]
            flat_idx = np.unravel_index(rng.integers(arr[i].size), arr[i].shape)
            idx = (i, *flat_idx)
            x0 = float(arr[idx])
            fds = []
            for h in (1e-6, 5e-7):
                sp = {k: v.astype(np.float64) for k, v in scene.items()}
                sp[name][idx] = x0 + h
                lp = _loss64(sp, cam, w, mode)
                sp[name][idx] = x0 - h
                lm = _loss64(sp, cam, w, mode)
                fds.append((lp - lm) / (2 * h))
            an = float(g[name][idx])
            scale = max(abs(fds[0]), abs(an), 1e-9)
            if abs(fds[0] - fds[1]) > 1e-3 * scale:
                skipped += 1  # a threshold (alpha 1/255, T 1e-4, rect) was crossed
                continue
            assert abs(fds[0] - an) <= 1e-4 * scale + 1e-10, (name, idx, fds, an)
            ok += 1
    assert ok >= 60 and skipped <= 8
