"""Model-file and camera I/O (fastergs_amd.plyio) against the file formats.

The PLY fixtures are written here byte by byte with struct, independently of
save_ply, in the vertex layout of 3DGS/scene/gaussian_model.py:177-208; the
camera checks use what the projection must give (closed forms), not the
loader's own formulas.
"""
import json
import math
import struct

import numpy as np
import pytest

from fastergs_amd import plyio, scenes


def _write_ply(path, names, rows, fmt="binary_little_endian", comment=True, ptype="float"):
    head = ["ply", f"format {fmt} 1.0"]
    if comment:
        head.append("comment written by a test")
    head.append(f"element vertex {len(rows)}")
    head += [f"property {ptype} {n}" for n in names]
    head.append("end_header")
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        if fmt == "ascii":
            for r in rows:
                f.write((" ".join(repr(float(x)) for x in r) + "\n").encode())
        else:
            e = "<" if fmt == "binary_little_endian" else ">"
            c = "f" if ptype == "float" else "d"
            for r in rows:
                f.write(struct.pack(e + c * len(r), *r))


def _names(K):
    return (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
            + [f"f_rest_{i}" for i in range(3 * K)] + ["opacity"]
            + [f"scale_{i}" for i in range(3)] + [f"rot_{i}" for i in range(4)])


def _row(g, K):
    """Raw values of Gaussian g, each exactly representable in float32."""
    xyz = [g + 0.5, -g - 0.25, 2.0 + g]
    dc = [0.125 * g, -0.5, 0.75]
    rest = [100.0 * g + i for i in range(3 * K)]
    return xyz + [0.0, 0.0, 0.0] + dc + rest + [0.5 * g - 1.0] + [-2.0, -3.0, -4.5 + g] \
        + [2.0, 0.0, -1.0, 2.0 * g]


@pytest.mark.parametrize("fmt", ["binary_little_endian", "binary_big_endian", "ascii"])
def test_load_raw_and_activated(tmp_path, fmt):
    K = 3  # degree 1
    rows = [_row(g, K) for g in range(3)]
    path = str(tmp_path / "m.ply")
    _write_ply(path, _names(K), rows, fmt=fmt)
    raw = plyio.load_ply(path, activated=False)
    act = plyio.load_ply(path)
    assert raw["sh_degree"] == 1 and act["sh_degree"] == 1
    for g in range(3):
        r = rows[g]
        assert raw["means3D"][g].tolist() == r[0:3]
        assert raw["dc"][g, 0].tolist() == r[6:9]
        # channel-major storage: f_rest[c*K + k] is coefficient k of channel c
        for c in range(3):
            for k in range(K):
                assert raw["shs"][g, k, c] == r[9 + c * K + k]
        assert raw["opacities"][g, 0] == r[9 + 3 * K]
        assert raw["scales"][g].tolist() == r[10 + 3 * K:13 + 3 * K]
        assert raw["rotations"][g].tolist() == r[13 + 3 * K:17 + 3 * K]
        assert act["opacities"][g, 0] == pytest.approx(1 / (1 + math.exp(-r[9 + 3 * K])), rel=1e-6)
        assert np.allclose(act["scales"][g], np.exp(r[10 + 3 * K:13 + 3 * K]), rtol=1e-6)
        q = np.array(r[13 + 3 * K:], dtype=np.float64)
        assert np.allclose(act["rotations"][g], q / np.linalg.norm(q), rtol=1e-6)
        assert act["shs"][g].tolist() == raw["shs"][g].tolist()


def test_property_order_and_extra_properties(tmp_path):
    """Properties are found by name: shuffled order, an extra property and
    f_rest indices past 9 (sorted numerically, not as strings) load alike."""
    K = 15  # degree 3: f_rest_0..44
    names = _names(K)
    rows = [_row(g, K) for g in range(2)]
    perm = np.random.default_rng(0).permutation(len(names))
    names_s = [names[i] for i in perm] + ["filter_3D"]
    rows_s = [[r[i] for i in perm] + [7.0] for r in rows]
    a, b = str(tmp_path / "a.ply"), str(tmp_path / "b.ply")
    _write_ply(a, names, rows)
    _write_ply(b, names_s, rows_s, ptype="double")
    A, B = plyio.load_ply(a, activated=False), plyio.load_ply(b, activated=False)
    assert A["sh_degree"] == 3
    for k in ("means3D", "opacities", "scales", "rotations", "dc", "shs"):
        assert np.array_equal(A[k], B[k]), k
    assert A["shs"][1, 14, 2] == rows[1][9 + 2 * K + 14]


def test_degree_zero_and_bad_counts(tmp_path):
    p = str(tmp_path / "d0.ply")
    _write_ply(p, _names(0), [_row(0, 0)])
    m = plyio.load_ply(p)
    assert m["sh_degree"] == 0 and m["shs"].shape == (1, 0, 3)
    bad = [n for n in _names(3) if n != "f_rest_8"]
    q = str(tmp_path / "bad.ply")
    _write_ply(q, bad, [_row(0, 3)[:-1]])
    with pytest.raises(ValueError):
        plyio.load_ply(q)
    t = str(tmp_path / "trunc.ply")
    _write_ply(t, _names(0), [_row(0, 0), _row(1, 0)])
    data = open(t, "rb").read()
    open(t, "wb").write(data[:-3])
    with pytest.raises(ValueError):
        plyio.load_ply(t)


def test_save_load_round_trip(tmp_path):
    scene = scenes.cluster_scene(500, seed=3)
    p = str(tmp_path / "out" / "point_cloud.ply")
    plyio.save_ply(p, scene, activated=False)
    back = plyio.load_ply(p, activated=False)
    for k in ("means3D", "opacities", "scales", "rotations", "dc", "shs"):
        assert np.array_equal(back[k], scene[k]), k
    plyio.save_ply(p, scene)  # activated: stored as logit / log
    act = plyio.load_ply(p)
    assert np.allclose(act["opacities"], scene["opacities"], rtol=1e-5, atol=1e-7)
    assert np.allclose(act["scales"], scene["scales"], rtol=1e-5)
    assert np.allclose(act["rotations"], scene["rotations"], atol=1e-6)


def _project(cam, x):
    """Pixel coordinates of world point x with the rasterizer's conventions
    (column-major matrices, ndc -> ((ndc + 1) * W - 1) / 2)."""
    h = np.append(np.asarray(x, dtype=np.float64), 1.0) @ cam.projmatrix.astype(np.float64)
    ndc = h[:3] / h[3]
    return ((ndc[0] + 1) * cam.width - 1) / 2, ((ndc[1] + 1) * cam.height - 1) / 2, h[3]


def test_camera_json_closed_form(tmp_path):
    """A camera at (0, 0, -4) looking along +z (camera-to-world rotation =
    identity), focal 1000 px on a 1280x840 image: the origin lands on the
    image centre at depth 4, a point 1 unit to the right lands 1000/4 px
    right of it, one unit down (+y) 250 px lower."""
    e = {"id": 0, "img_name": "a", "width": 1280, "height": 840, "position": [0.0, 0.0, -4.0],
         "rotation": [[1, 0, 0], [0, 1, 0], [0, 0, 1]], "fx": 1000.0, "fy": 1000.0}
    cam = plyio.camera_from_json(e)
    assert cam.tanfovx == pytest.approx(640 / 1000) and cam.tanfovy == pytest.approx(420 / 1000)
    assert np.allclose(cam.campos, [0, 0, -4], atol=1e-6)
    u, v, w = _project(cam, [0, 0, 0])
    assert (u, v) == pytest.approx((639.5, 419.5), abs=1e-3) and w == pytest.approx(4.0, rel=1e-6)
    u1, v1, _ = _project(cam, [1, 0, 0])
    u2, v2, _ = _project(cam, [0, 1, 0])
    assert u1 - u == pytest.approx(250.0, abs=1e-2) and v1 == pytest.approx(v, abs=1e-3)
    assert v2 - v == pytest.approx(250.0, abs=1e-2) and u2 == pytest.approx(u, abs=1e-3)
    # a camera turned 90 degrees about y (looking along +x from (-4, 0, 0))
    rot = [[0, 0, 1], [0, 1, 0], [-1, 0, 0]]
    cam2 = plyio.camera_from_json(dict(e, position=[-4.0, 0.0, 0.0], rotation=rot))
    u, v, w = _project(cam2, [0, 0, 0])
    assert (u, v) == pytest.approx((639.5, 419.5), abs=1e-3) and w == pytest.approx(4.0, rel=1e-6)


def test_cameras_json_round_trip(tmp_path):
    cams = scenes.sphere_cameras(6, seed=2)
    p = str(tmp_path / "cameras.json")
    plyio.save_cameras_json(p, cams)
    entries = json.load(open(p))
    assert [e["id"] for e in entries] == list(range(6))
    back = plyio.load_cameras_json(p)
    for a, b in zip(cams, back):
        assert (a.width, a.height) == (b.width, b.height)
        assert a.tanfovx == pytest.approx(b.tanfovx, rel=1e-6)
        assert a.tanfovy == pytest.approx(b.tanfovy, rel=1e-6)
        assert np.allclose(a.viewmatrix, b.viewmatrix, atol=1e-5)
        assert np.allclose(a.projmatrix, b.projmatrix, rtol=1e-5, atol=1e-5)
        assert np.allclose(a.campos, b.campos, atol=1e-5)
