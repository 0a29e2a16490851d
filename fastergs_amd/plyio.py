"""Trained-scene and camera I/O for replaying real scenes through the rasterizer.

PLY: the vertex layout 3DGS trainers write and read
(3DGS/scene/gaussian_model.py:177-189 attribute list, :191-208 save_ply,
:215-256 load_ply; camera centre as in 3DGS/scene/cameras.py:54-57): one ``vertex`` element of float32 properties
x, y, z, nx, ny, nz, f_dc_0..2, f_rest_0..(3K-1), opacity, scale_0..2,
rot_0..3, where K = (degree+1)^2 - 1 and f_rest is stored channel-major
(all K coefficients of red, then green, then blue). Stored values are the
raw parameters: opacity before the sigmoid, log scales, unnormalised
quaternions (w, x, y, z). ``load_ply`` returns the activated attributes the
rasterizer takes (3DGS/gaussian_renderer/render_faster.py:100-105), or the
raw ones with ``activated=False``.

Cameras: the ``cameras.json`` entries 3DGS writes next to the model
(3DGS/utils/camera_utils.py:62-82): ``position`` is the camera centre,
``rotation`` the camera-to-world rotation (3x3 rows), ``fx``/``fy`` the focal
lengths in pixels. The matrices follow 3DGS/utils/graphics_utils.py:38-77
(znear 0.01, zfar 100).
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from .scenes import Camera, projection_matrix

_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1",
    "short": "i2", "int16": "i2", "ushort": "u2", "uint16": "u2",
    "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


def _read_header(f):
    """Parse a PLY header; returns (format, [(element, count, [(name, dtype)])])."""
    if f.readline().strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt = None
    elements = []
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header has no end_header")
        words = line.decode("ascii").split()
        if not words or words[0] in ("comment", "obj_info"):
            continue
        if words[0] == "end_header":
            break
        if words[0] == "format":
            fmt = words[1]
        elif words[0] == "element":
            elements.append((words[1], int(words[2]), []))
        elif words[0] == "property":
            if not elements:
                raise ValueError("PLY property before any element")
            if words[1] == "list":
                raise ValueError("PLY list properties are not supported")
            if words[1] not in _PLY_TYPES:
                raise ValueError(f"unknown PLY property type {words[1]!r}")
            elements[-1][2].append((words[2], _PLY_TYPES[words[1]]))
    if fmt not in ("binary_little_endian", "binary_big_endian", "ascii"):
        raise ValueError(f"unsupported PLY format {fmt!r}")
    return fmt, elements


def read_ply_vertices(path: str) -> np.ndarray:
    """The ``vertex`` element of a PLY file as a numpy structured array."""
    with open(path, "rb") as f:
        fmt, elements = _read_header(f)
        endian = ">" if fmt == "binary_big_endian" else "<"
        out = None
        for name, count, props in elements:
            dt = np.dtype([(p, endian + t) for p, t in props])
            if fmt == "ascii":
                rows = [f.readline().split() for _ in range(count)]
                arr = np.empty(count, dtype=dt)
                for j, (p, _) in enumerate(props):
                    arr[p] = [r[j] for r in rows]
            else:
                raw = f.read(dt.itemsize * count)
                if len(raw) != dt.itemsize * count:
                    raise ValueError(f"PLY element {name!r} is truncated")
                arr = np.frombuffer(raw, dtype=dt, count=count)
            if name == "vertex":
                out = arr
                break
    if out is None:
        raise ValueError("PLY file has no vertex element")
    return out


def _indexed(names, prefix):
    sel = [n for n in names if n.startswith(prefix)]
    return sorted(sel, key=lambda n: int(n.split("_")[-1]))


def load_ply(path: str, activated: bool = True) -> dict:
    """A 3DGS model file as rasterizer inputs (numpy float32): means3D (P,3),
    opacities (P,1), scales (P,3), rotations (P,4), dc (P,1,3), shs (P,K,3),
    and ``sh_degree``. ``activated`` applies sigmoid / exp / normalisation as
    the trainers' getters do."""
    v = read_ply_vertices(path)
    names = v.dtype.names
    P = v.shape[0]

    def cols(keys):
        return np.stack([np.asarray(v[k], dtype=np.float64) for k in keys], axis=1)

    xyz = cols(["x", "y", "z"])
    dc = cols(["f_dc_0", "f_dc_1", "f_dc_2"]).reshape(P, 1, 3)
    rest_names = _indexed(names, "f_rest_")
    if len(rest_names) % 3:
        raise ValueError("f_rest_* count is not a multiple of 3")
    K = len(rest_names) // 3
    degree = round(math.sqrt(K + 1)) - 1
    if (degree + 1) ** 2 - 1 != K or degree > 3:
        raise ValueError(f"{len(rest_names)} f_rest_* properties match no SH degree <= 3")
    # stored channel-major: (P, 3, K) -> (P, K, 3)
    shs = (cols(rest_names).reshape(P, 3, K).transpose(0, 2, 1) if K
           else np.zeros((P, 0, 3)))
    opacity = cols(["opacity"])
    scales = cols(_indexed(names, "scale_"))
    rots = cols(_indexed(names, "rot"))
    if scales.shape[1] != 3 or rots.shape[1] != 4:
        raise ValueError("expected scale_0..2 and rot_0..3")
    if activated:
        opacity = 1.0 / (1.0 + np.exp(-opacity))
        scales = np.exp(scales)
        n = np.linalg.norm(rots, axis=1, keepdims=True)
        rots = rots / np.maximum(n, 1e-12)  # torch.nn.functional.normalize's eps
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    return {"means3D": f32(xyz), "opacities": f32(opacity), "scales": f32(scales),
            "rotations": f32(rots), "dc": f32(dc), "shs": f32(shs), "sh_degree": degree}


def save_ply(path: str, scene: dict, activated: bool = True) -> None:
    """Write rasterizer attributes as a 3DGS model file (binary little endian).
    With ``activated`` the opacities and scales are mapped back to the stored
    raw parameters (logit, log); rotations are stored as given."""
    xyz = np.asarray(scene["means3D"], dtype=np.float64)
    P = xyz.shape[0]
    dc = np.asarray(scene["dc"], dtype=np.float64).reshape(P, 3)
    shs = np.asarray(scene.get("shs", np.zeros((P, 0, 3))), dtype=np.float64)
    K = shs.shape[1]
    rest = shs.transpose(0, 2, 1).reshape(P, 3 * K)
    op = np.asarray(scene["opacities"], dtype=np.float64).reshape(P, 1)
    sc = np.asarray(scene["scales"], dtype=np.float64).reshape(P, 3)
    if activated:
        op = np.log(op / (1.0 - op))
        sc = np.log(sc)
    rot = np.asarray(scene["rotations"], dtype=np.float64).reshape(P, 4)
    names = (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
             + [f"f_rest_{i}" for i in range(3 * K)] + ["opacity"]
             + [f"scale_{i}" for i in range(3)] + [f"rot_{i}" for i in range(4)])
    data = np.concatenate([xyz, np.zeros_like(xyz), dc, rest, op, sc, rot], axis=1)
    header = ["ply", "format binary_little_endian 1.0", f"element vertex {P}"]
    header += [f"property float {n}" for n in names] + ["end_header"]
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n").encode("ascii"))
        f.write(np.ascontiguousarray(data, dtype="<f4").tobytes())


def camera_from_json(entry: dict, znear: float = 0.01, zfar: float = 100.0) -> Camera:
    """One cameras.json entry as a rasterizer camera."""
    W, H = int(entry["width"]), int(entry["height"])
    fovx = 2.0 * math.atan(W / (2.0 * float(entry["fx"])))
    fovy = 2.0 * math.atan(H / (2.0 * float(entry["fy"])))
    c2w_rot = np.asarray(entry["rotation"], dtype=np.float64).reshape(3, 3)
    pos = np.asarray(entry["position"], dtype=np.float64).reshape(3)
    Rv = c2w_rot.T  # world -> view rotation
    V = np.eye(4)
    V[:3, :3] = Rv
    V[:3, 3] = -Rv @ pos
    world_view = V.T.astype(np.float32)
    P = projection_matrix(znear, zfar, fovx, fovy)
    full_proj = (world_view.astype(np.float64) @ P.T).astype(np.float32)
    campos = np.linalg.inv(world_view.astype(np.float64))[3, :3].astype(np.float32)
    return Camera(W, H, math.tan(fovx / 2), math.tan(fovy / 2), world_view, full_proj, campos)


def camera_to_json(cam: Camera, cam_id: int = 0, name: str = "") -> dict:
    """The cameras.json entry of a rasterizer camera."""
    V = np.asarray(cam.viewmatrix, dtype=np.float64).T  # world -> view
    Rv = V[:3, :3]
    pos = -Rv.T @ V[:3, 3]
    return {"id": cam_id, "img_name": name, "width": cam.width, "height": cam.height,
            "position": pos.tolist(), "rotation": Rv.T.tolist(),
            "fy": cam.height / (2.0 * cam.tanfovy), "fx": cam.width / (2.0 * cam.tanfovx)}


def load_cameras_json(path: str) -> list:
    with open(path) as f:
        entries = json.load(f)
    return [camera_from_json(e) for e in entries]


def save_cameras_json(path: str, cams) -> None:
    with open(path, "w") as f:
        json.dump([camera_to_json(c, i, f"view_{i:05d}") for i, c in enumerate(cams)], f)
