"""Pins of the oracle's render_simp statistics on the CPU.

* One Gaussian: its weight sum is the sum over the pixels of its tiles of
  alpha = min(0.99, 2^q) (q from the splat record, alpha >= 1/255, float64
  here), its count the number of such pixels, and it is every such pixel's
  maximum.
* Many Gaussians: every in-image pixel that met a splat credits exactly one
  maximum, so the maxima sum to the number of such pixels; and the weights
  of a pixel sum to 1 - T, so the total weight is bounded by the pixel count.
"""
import numpy as np

from fastergs_amd import scenes
from oracle.pyoracle import Oracle


def _single():
    s = scenes.uniform_scene(1, seed=0)
    s["means3D"][:] = [[0.05, -0.02, 0.0]]
    s["scales"][:] = [[0.05, 0.03, 0.04]]
    s["opacities"][:] = [[0.8]]
    return s


def test_single_gaussian_closed_form():
    scene = _single()
    cam = scenes.look_at_camera((0.0, 0.0, -4.0), 200, 150)
    orc = Oracle(scene, cam, bucket=32, mode=0).forward()
    w, c, m = orc.render_simp()
    sp = orc.get("splat").reshape(-1, 9)[0].astype(np.float64)
    rng = orc.get("ranges").reshape(-1, 2)
    gx = (cam.width + 15) // 16
    total, count = 0.0, 0
    for t in np.nonzero(rng[:, 1] > rng[:, 0])[0]:
        tx, ty = t % gx, t // gx
        for py in range(ty * 16, min(ty * 16 + 16, cam.height)):
            for px in range(tx * 16, min(tx * 16 + 16, cam.width)):
                dx, dy = sp[0] - px, sp[1] - py
                q = sp[5] + sp[2] * dx * dx + sp[4] * dy * dy + sp[3] * dx * dy
                a = min(0.99, 2.0 ** q)
                if a >= 1.0 / 255.0:
                    total += a
                    count += 1
    assert count > 50
    assert abs(int(c[0]) - count) <= 2  # float32 vs float64 at the 1/255 edge
    assert abs(w[0] - total) <= 1e-5 * total + 2.0 / 255.0
    assert m[0] == c[0]


def test_many_gaussians_invariants():
    scene = scenes.cluster_scene(3000, seed=5, clusters=8, radius=1.0)
    cam = scenes.look_at_camera((0.3, 0.1, -4.0), 150, 110)
    orc = Oracle(scene, cam, bucket=32, mode=0).forward()
    w, c, m = orc.render_simp()
    assert m.sum() <= cam.width * cam.height
    assert (m <= c).all()                       # a maximum is a contributing pixel
    assert w.sum() <= cam.width * cam.height    # per pixel the weights sum to 1 - T
    assert (c > 0).sum() > 100 and m.sum() > 0.5 * cam.width * cam.height
