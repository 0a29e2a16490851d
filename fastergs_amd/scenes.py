"""Seeded synthetic scenes and pinhole cameras for the rasterizer workloads.

The camera matrices follow the 3DGS conventions the FasterGS callers use
(3DGS/scene/cameras.py:48-57, 3DGS/utils/graphics_utils.py:38-71):
view space is +z forward / +y down, ``viewmatrix`` is the world->view matrix
stored transposed (so the kernels read it column-major), ``projmatrix`` is
``viewmatrix @ P^T`` with znear 0.01 and zfar 100, and ``campos`` is the
camera centre. FoV is applied to the vertical axis; fovx follows from the
aspect ratio. Attributes are the *activated* values the callers pass
(3DGS/gaussian_renderer/render_faster.py:100-105): opacity after sigmoid,
scales after exp, unit quaternions stored (w, x, y, z).

Workloads (sizes as stated in the project's benchmark table):
  small    100k Gaussians, uniform ball r=2, log-scale N(-4,0.7), 800x600
  heavy    1M Gaussians, 64-cluster mixture (sigma 0.3) in r=3 ball,
           98% log-scale N(-4.5,1.0) + 2% N(-1,0.5), 1280x840
  dense    3M Gaussians, 256 clusters, opacity sigmoid(N(1,1)), 1920x1080,
           120-frame orbit at r=3.5
  multiview 2M heavy-distributed Gaussians, 16 views of uneven resolution
  large    6M heavy-distributed Gaussians, 1237x822, 200-view trajectory
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

SH_REST = 15  # degree-3 rest coefficients


@dataclass
class Camera:
    width: int
    height: int
    tanfovx: float
    tanfovy: float
    viewmatrix: np.ndarray  # (4,4) float32, transposed world->view
    projmatrix: np.ndarray  # (4,4) float32, viewmatrix @ P^T
    campos: np.ndarray  # (3,) float32


def projection_matrix(znear: float, zfar: float, fovx: float, fovy: float) -> np.ndarray:
    """Perspective matrix of 3DGS/utils/graphics_utils.py:51-71 (not transposed)."""
    tan_y = math.tan(fovy / 2)
    tan_x = math.tan(fovx / 2)
    top, right = tan_y * znear, tan_x * znear
    P = np.zeros((4, 4), dtype=np.float64)
    P[0, 0] = 2.0 * znear / (2 * right)
    P[1, 1] = 2.0 * znear / (2 * top)
    P[3, 2] = 1.0
    P[2, 2] = zfar / (zfar - znear)
    P[2, 3] = -(zfar * znear) / (zfar - znear)
    return P


def look_at_camera(pos, width: int, height: int, fovy_deg: float = 60.0, target=(0.0, 0.0, 0.0),
                   up=(0.0, 1.0, 0.0)) -> Camera:
    pos = np.asarray(pos, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - pos
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, dtype=np.float64))
    if np.linalg.norm(r) < 1e-6:
        r = np.cross(f, np.array([1.0, 0.0, 0.0]))
    r /= np.linalg.norm(r)
    d = np.cross(f, r)  # image y points down
    Rv = np.stack([r, d, f], axis=0)  # world -> view rotation (rows)
    V = np.eye(4)
    V[:3, :3] = Rv
    V[:3, 3] = -Rv @ pos
    fovy = math.radians(fovy_deg)
    fovx = 2.0 * math.atan(math.tan(fovy / 2) * width / height)
    P = projection_matrix(0.01, 100.0, fovx, fovy)
    world_view = V.T
    full_proj = world_view @ P.T
    return Camera(width, height, math.tan(fovx / 2), math.tan(fovy / 2),
                  world_view.astype(np.float32), full_proj.astype(np.float32),
                  pos.astype(np.float32))


def _unit_quats(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.astype(np.float32)


def _ball(rng, n, radius):
    v = rng.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    r = radius * rng.random(n) ** (1.0 / 3.0)
    return v * r[:, None]


def _sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def _attributes(rng, n, means, log_scale, opacity_mu=0.0):
    return {
        "means3D": means.astype(np.float32),
        "opacities": _sigmoid(rng.normal(opacity_mu, 1.0, (n, 1))).astype(np.float32),
        "scales": np.exp(log_scale).astype(np.float32),
        "rotations": _unit_quats(rng, n),
        "dc": rng.normal(0.0, 0.5, (n, 1, 3)).astype(np.float32),
        "shs": rng.normal(0.0, 0.05, (n, SH_REST, 3)).astype(np.float32),
    }


def uniform_scene(n: int, seed: int = 0, radius: float = 2.0, log_mu: float = -4.0,
                  log_sigma: float = 0.7) -> dict:
    rng = np.random.default_rng(seed)
    means = _ball(rng, n, radius)
    log_scale = rng.normal(log_mu, log_sigma, (n, 3))
    return _attributes(rng, n, means, log_scale)


def cluster_scene(n: int, seed: int = 0, clusters: int = 64, radius: float = 3.0,
                  sigma: float = 0.3, big_frac: float = 0.02, opacity_mu: float = 0.0) -> dict:
    rng = np.random.default_rng(seed)
    centres = _ball(rng, clusters, radius)
    which = rng.integers(0, clusters, n)
    means = centres[which] + rng.normal(0.0, sigma, (n, 3))
    big = rng.random(n) < big_frac
    log_scale = np.where(big[:, None], rng.normal(-1.0, 0.5, (n, 3)), rng.normal(-4.5, 1.0, (n, 3)))
    return _attributes(rng, n, means, log_scale, opacity_mu)


WORKLOADS = {
    "small": dict(P=100_000, W=800, H=600),
    "heavy": dict(P=1_000_000, W=1280, H=840),
    "dense": dict(P=3_000_000, W=1920, H=1080),
    "multiview": dict(P=2_000_000, W=1280, H=840),
    "large": dict(P=6_000_000, W=1237, H=822),
}


def workload(name: str, seed: int = 0, n: int | None = None):
    """Scene plus its first camera for a named workload (optionally resized to n Gaussians)."""
    spec = WORKLOADS[name]
    P = spec["P"] if n is None else n
    if name == "small":
        scene = uniform_scene(P, seed)
        cam = look_at_camera((0.0, 0.0, -4.0), spec["W"], spec["H"])
    elif name == "dense":
        scene = cluster_scene(P, seed, clusters=256, opacity_mu=1.0)
        cam = orbit_cameras(1, spec["W"], spec["H"], radius=3.5)[0]
    else:
        scene = cluster_scene(P, seed)
        cam = look_at_camera((0.0, 0.0, -4.0), spec["W"], spec["H"])
    return scene, cam


def orbit_cameras(frames: int, width: int, height: int, radius: float = 3.5, height_y: float = 0.5):
    cams = []
    for k in range(frames):
        a = 2.0 * math.pi * k / frames
        pos = (radius * math.sin(a), height_y, -radius * math.cos(a))
        cams.append(look_at_camera(pos, width, height))
    return cams


def sphere_cameras(n: int, seed: int = 0, rmin: float = 3.0, rmax: float = 5.0,
                   sizes=((640, 480), (960, 720), (1280, 840), (1920, 1080))):
    rng = np.random.default_rng(seed + 1000)
    cams = []
    for _ in range(n):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        r = rng.uniform(rmin, rmax)
        W, H = sizes[rng.integers(0, len(sizes))]
        cams.append(look_at_camera(d * r, int(W), int(H)))
    return cams


def trajectory_cameras(n: int, width: int, height: int, seed: int = 0, rmin: float = 3.0,
                       rmax: float = 5.0, step_deg: float = 4.0):
    """A seeded random camera trajectory looking at the origin: the viewing
    direction takes a random-walk step of about ``step_deg`` degrees per view
    (a rotation about a random axis normal to it) and the distance drifts
    smoothly between rmin and rmax."""
    rng = np.random.default_rng(seed + 2000)
    d = rng.standard_normal(3)
    d /= np.linalg.norm(d)
    phase, dphase = rng.uniform(0.0, 2.0 * math.pi), rng.uniform(0.03, 0.08)
    cams = []
    for _ in range(n):
        r = rmin + (rmax - rmin) * 0.5 * (1.0 + math.sin(phase))
        cams.append(look_at_camera(d * r, width, height))
        axis = np.cross(d, rng.standard_normal(3))
        axis /= np.linalg.norm(axis)
        a = math.radians(step_deg) * rng.uniform(0.5, 1.5)
        d = d * math.cos(a) + np.cross(axis, d) * math.sin(a)  # Rodrigues, axis normal to d
        d /= np.linalg.norm(d)
        phase += dphase
    return cams


def fd_scene(n: int = 1000, seed: int = 0) -> tuple[dict, Camera]:
    """Small in-frustum scene for finite-difference gradient checks: 128x128,
    Gaussians of a few pixels, opacities away from the 1/255 cut-off."""
    rng = np.random.default_rng(seed)
    means = rng.uniform(-0.9, 0.9, (n, 3)) * np.array([1.0, 1.0, 0.5])
    log_scale = rng.normal(-3.3, 0.3, (n, 3))
    scene = _attributes(rng, n, means, log_scale)
    scene["opacities"] = rng.uniform(0.05, 0.6, (n, 1)).astype(np.float32)
    cam = look_at_camera((0.0, 0.0, -4.0), 128, 128)
    return scene, cam


def target_image(width: int, height: int, seed: int = 0) -> np.ndarray:
    return np.random.default_rng(seed + 7).random((3, height, width)).astype(np.float32)
