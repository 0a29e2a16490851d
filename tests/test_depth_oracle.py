"""Pins of the oracle's render_depth on the CPU.

* Geometry in float64 against closed forms: the 4x4 inverse against
  numpy.linalg.inv; each pixel's ray against the view-space point
  (ndc_x tanfovx, ndc_y tanfovy, 1) of the camera (the normalised depth the
  reference uses maps to view depth 1 for a 0.01 / 100 frustum) and against
  numpy's inverse of the stored projection; the
  ellipsoid quadratic against x^T A x with A = (9 Sigma)^-1 and
  Sigma = Rq S^2 Rq^T, Rq from scipy's quaternion rotation.
* One Gaussian: every pixel it covers reports index 0, T = 1 - alpha, the
  depth of the midpoint of its 3-sigma chord along the pixel's ray and the
  point there (where the ray is closest to the mean in the Mahalanobis
  metric); uncovered pixels keep zeros and T = 1.
* A small scene against a float64 walk of the lists (alphas from the splat
  records, geometry from the closed forms above) that applies the selection
  rule: stop when T (1 - alpha) < 1e-4, skip hits behind the camera, keep
  the first largest weight alpha T.
"""
import numpy as np
from scipy.spatial.transform import Rotation

from fastergs_amd import scenes
from oracle.pyoracle import Oracle, _ptr, lib


def _view_ray(cam, px, py):
    """Unit direction and length of the reference ray, from the camera's
    view matrix and field of view (float64)."""
    nx = (2.0 * px + 1.0) / cam.width - 1.0
    ny = (2.0 * py + 1.0) / cam.height - 1.0
    pv = np.array([nx * 1.0000001 * cam.tanfovx, ny * 1.0000001 * cam.tanfovy, 1.0])
    V = cam.viewmatrix.astype(np.float64).T          # world -> view
    world = np.linalg.solve(V[:3, :3], pv - V[:3, 3])
    d = world - cam.campos.astype(np.float64)
    return d / np.linalg.norm(d), np.linalg.norm(d)


def _stored_ray(cam, px, py):
    """The reference's construction on the stored projection, in float64."""
    nx = (2.0 * px + 1.0) / cam.width - 1.0
    ny = (2.0 * py + 1.0) / cam.height - 1.0
    h = np.array([nx * 1.0000001, ny * 1.0000001, (100 - 100 * 0.01) / (100 - 0.01), 1.0])
    d = (np.linalg.inv(cam.projmatrix.astype(np.float64).T) @ h)[:3] - cam.campos
    return d / np.linalg.norm(d), np.linalg.norm(d)


def _quadratic(mean, scale, q, origin, ndir):
    Rq = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()
    A = np.linalg.inv(9.0 * Rq @ np.diag(np.asarray(scale) ** 2) @ Rq.T)
    d0 = np.asarray(origin) - np.asarray(mean)
    a = ndir @ A @ ndir
    b = 2.0 * (ndir @ A @ d0)
    c = d0 @ A @ d0 - 1.0
    return -b / (2.0 * a), b * b - 4.0 * a * c, A, b * b + abs(4.0 * a * c)


def _f32_ray(cam, px, py):
    """The float32 ray the render path uses (its construction is pinned in
    float64 above; in float32 it carries ~2e-5 of cancellation error)."""
    proj = np.ascontiguousarray(cam.projmatrix.reshape(16), dtype=np.float32)
    cp = np.ascontiguousarray(cam.campos, dtype=np.float32)
    out = np.zeros(4, np.float32)
    lib().orc_depth_ray_f32(_ptr(proj), _ptr(cp), int(px), int(py), cam.width, cam.height,
                            _ptr(out))
    return out[:3].astype(np.float64), float(out[3])


def test_inverse4_matches_numpy():
    L = lib()
    rng = np.random.default_rng(0)
    for _ in range(20):
        m = rng.normal(size=(4, 4)) + 3.0 * np.eye(4)
        flat = np.ascontiguousarray(m.reshape(16))
        inv = np.zeros(16)
        L.orc_inverse4_f64(_ptr(flat), _ptr(inv))
        np.testing.assert_allclose(inv.reshape(4, 4), np.linalg.inv(m), rtol=1e-10, atol=1e-12)
    sing = np.ascontiguousarray(np.ones(16))
    inv = np.full(16, 7.0)
    L.orc_inverse4_f64(_ptr(sing), _ptr(inv))
    assert (inv == 0.0).all()


def test_pixel_ray_closed_form():
    L = lib()
    for pos, (W, H) in [((0.0, 0.0, -4.0), (200, 150)), ((1.5, 0.7, 3.0), (97, 131))]:
        cam = scenes.look_at_camera(pos, W, H, target=(0.2, -0.1, 0.3))
        proj = np.ascontiguousarray(cam.projmatrix.astype(np.float64).reshape(16))
        cp = np.ascontiguousarray(cam.campos.astype(np.float64))
        for px, py in [(0, 0), (W - 1, 0), (0, H - 1), (W - 1, H - 1), (W // 2, H // 3)]:
            out = np.zeros(4)
            L.orc_depth_ray_f64(_ptr(proj), _ptr(cp), px, py, W, H, _ptr(out))
            # the same construction through numpy's inverse of the stored matrix
            nx = (2.0 * px + 1.0) / W - 1.0
            ny = (2.0 * py + 1.0) / H - 1.0
            h = np.array([nx * 1.0000001, ny * 1.0000001, (100 - 100 * 0.01) / (100 - 0.01), 1.0])
            d = (np.linalg.inv(proj.reshape(4, 4).T) @ h)[:3] - cp
            np.testing.assert_allclose(out[:3], d / np.linalg.norm(d), rtol=1e-9, atol=1e-12)
            np.testing.assert_allclose(out[3], np.linalg.norm(d), rtol=1e-9)
            # and the geometric ray of the camera: the float32 rounding of the
            # projection's depth row is amplified ~100x by the cancellation
            # that maps the normalised depth back to view depth 1
            nd, ln = _view_ray(cam, px, py)
            np.testing.assert_allclose(out[:3], nd, atol=1e-4)
            np.testing.assert_allclose(out[3], ln, rtol=1e-4)


def test_ellipsoid_quadratic_closed_form():
    L = lib()
    rng = np.random.default_rng(1)
    for _ in range(50):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        mean = rng.normal(size=3)
        scale = np.exp(rng.normal(size=3) - 1.0)
        origin = rng.normal(size=3) * 3.0
        nd = rng.normal(size=3)
        nd /= np.linalg.norm(nd)
        out = np.zeros(2)
        args = [np.ascontiguousarray(v, dtype=np.float64) for v in (mean, scale, q, origin, nd)]
        L.orc_ellipsoid_f64(*[_ptr(v) for v in args], _ptr(out))
        t, disc, _, _ = _quadratic(mean, scale, q, origin, nd)
        np.testing.assert_allclose(out[0], t, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(out[1], disc, rtol=1e-8, atol=1e-10)


def _single():
    s = scenes.uniform_scene(1, seed=0)
    s["means3D"][:] = [[2.6, 1.9, 0.0]]   # towards a corner: |ray| ~ 1.2
    s["scales"][:] = [[0.05, 0.03, 0.04]]
    s["opacities"][:] = [[0.8]]
    return s


def test_single_gaussian_depth_closed_form():
    scene = _single()
    cam = scenes.look_at_camera((0.3, 0.2, -4.0), 200, 150, target=(0.3, 0.2, 0.0))
    orc = Oracle(scene, cam, bucket=32, mode=0).forward()
    pts, depth, alpha, gidx, disc = orc.render_depth()
    tf = orc.get("tfinal").reshape(cam.height, cam.width)
    hit = tf < 1.0
    assert hit.sum() > 50
    assert (gidx == 0).all()
    np.testing.assert_allclose(alpha[hit], tf[hit], atol=2e-7)
    assert (alpha[~hit] == 1.0).all() and (depth[~hit] == 0.0).all() and (pts[:, ~hit] == 0).all()
    mean = scene["means3D"][0].astype(np.float64)
    for py, px in zip(*np.nonzero(hit)):
        nd, ln = _f32_ray(cam, px, py)
        nd64, ln64 = _stored_ray(cam, px, py)
        assert np.abs(nd - nd64).max() < 1e-4 and ln > 1.15
        t, dsc, A, mag = _quadratic(mean, scene["scales"][0], scene["rotations"][0], cam.campos, nd)
        assert abs(depth[py, px] - t / ln) <= 2e-6 * abs(t / ln)
        assert abs(disc[py, px] - dsc) <= 1e-6 * mag     # b^2 - 4ac cancels in float32
        p = pts[:, py, px].astype(np.float64)
        np.testing.assert_allclose(p, cam.campos + t * nd, atol=5e-6 * np.abs(p).max())
        # the point is where the ray is closest to the mean in the A metric
        assert abs(nd @ A @ (p - mean)) <= 1e-3 * np.sqrt(nd @ A @ nd)


def test_small_scene_against_list_walk():
    scene = scenes.uniform_scene(60, seed=3, radius=0.8, log_mu=-2.5)
    cam = scenes.look_at_camera((0.2, -0.3, -3.0), 64, 48)
    orc = Oracle(scene, cam, bucket=32, mode=0).forward()
    pts, depth, alpha, gidx, disc = orc.render_depth()
    sp = orc.get("splat").reshape(-1, 9).astype(np.float64)
    rng_ = orc.get("ranges").reshape(-1, 2)
    vals = orc.get("vals")
    gx = (cam.width + 15) // 16
    quad = {}
    mism, checked = 0, 0
    for py in range(cam.height):
        for px in range(cam.width):
            t_id = (py // 16) * gx + px // 16
            nd, ln = _f32_ray(cam, px, py)
            T, wmax, best = 1.0, 0.0, None
            for k in range(rng_[t_id, 0], rng_[t_id, 1]):
                g = int(vals[k])
                dx, dy = sp[g, 0] - px, sp[g, 1] - py
                a = min(0.99, 2.0 ** (sp[g, 5] + sp[g, 2] * dx * dx + sp[g, 4] * dy * dy
                                      + sp[g, 3] * dx * dy))
                if a < 1.0 / 255.0:
                    continue
                if T * (1.0 - a) < 1e-4:
                    break
                if (g, px, py) not in quad:
                    quad[(g, px, py)] = _quadratic(scene["means3D"][g], scene["scales"][g],
                                                   scene["rotations"][g], cam.campos, nd)
                t = quad[(g, px, py)][0]
                if t / ln < 0.0:
                    continue
                if wmax < a * T:
                    wmax, best = a * T, (g, t / ln)
                T *= 1.0 - a
            checked += 1
            if best is None:
                mism += int(gidx[py, px] != 0 or depth[py, px] != 0.0)
                continue
            if gidx[py, px] != best[0]:
                mism += 1
                continue
            assert abs(depth[py, px] - best[1]) <= 1e-5 * abs(best[1])
            assert abs(alpha[py, px] - T) <= 1e-5
    assert checked == cam.width * cam.height
    assert mism <= checked // 200, mism   # float32 vs float64 near ties and the 1/255 edge
    assert (gidx != 0).sum() > checked // 4
