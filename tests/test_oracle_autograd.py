"""Independent PyTorch FP64 restatement of the rasterizer forward, differentiated
by autograd, against the oracle's hand-written FP64 adjoint (no GPU).

The restatement is written from the mathematics (EWA splatting with the 3DGS
conventions, real SH basis, front-to-back alpha compositing); it shares no
code with fastergs_amd/csrc/fgs_math.h. The discrete structure -- which
(tile, Gaussian) pairs exist, their order and where each pixel stops -- is
taken from the oracle, so only the differentiable arithmetic is compared.
"""
import math

import numpy as np
import pytest
import torch

from fastergs_amd import scenes
from oracle.pyoracle import Oracle

C0 = 0.28209479177387814
C1 = 0.4886025119029199
C2 = [1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
      0.5462742152960396]
C3 = [-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
      -0.4570457994644658, 1.445305721320277, -0.5900435899266435]


def sh_rgb(dirs, dc, rest):
    x, y, z = dirs[:, 0:1], dirs[:, 1:2], dirs[:, 2:3]
    basis = [C0 * torch.ones_like(x), -C1 * y, C1 * z, -C1 * x,
             C2[0] * x * y, C2[1] * y * z, C2[2] * (2 * z * z - x * x - y * y), C2[3] * x * z,
             C2[4] * (x * x - y * y),
             C3[0] * y * (3 * x * x - y * y), C3[1] * x * y * z,
             C3[2] * y * (4 * z * z - x * x - y * y),
             C3[3] * z * (2 * z * z - 3 * x * x - 3 * y * y),
             C3[4] * x * (4 * z * z - x * x - y * y), C3[5] * z * (x * x - y * y),
             C3[6] * x * (x * x - 3 * y * y)]
    coefs = torch.cat([dc, rest], dim=1)  # (P,16,3)
    out = sum(b * coefs[:, k, :] for k, b in enumerate(basis)) + 0.5
    return torch.clamp(out, min=0.0)


def quat_to_rot(q):
    r, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return torch.stack([
        torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y)], -1),
        torch.stack([2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x)], -1),
        torch.stack([2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y)], -1),
    ], -2)


def render_torch(p, cam, orc, bg):
    V = torch.from_numpy(cam.viewmatrix.astype(np.float64))  # stored transposed
    Pm = torch.from_numpy(cam.projmatrix.astype(np.float64))
    campos = torch.from_numpy(cam.campos.astype(np.float64))
    P = p["means3D"].shape[0]
    ones = torch.ones(P, 1, dtype=torch.float64)
    ph = torch.cat([p["means3D"], ones], 1)
    t = (ph @ V)[:, :3]
    hom = ph @ Pm
    ndc = hom[:, :2] / (hom[:, 3:4] + 1e-7)
    W, H = cam.width, cam.height
    px = ((ndc[:, 0] + 1) * W - 1) * 0.5
    py = ((ndc[:, 1] + 1) * H - 1) * 0.5
    R = quat_to_rot(p["rotations"])
    S2 = torch.diag_embed(p["scales"] ** 2)
    Sigma = R @ S2 @ R.transpose(1, 2)
    fx = W / (2 * float(np.float32(cam.tanfovx)))
    fy = H / (2 * float(np.float32(cam.tanfovy)))
    z = t[:, 2]
    J = torch.zeros(P, 2, 3, dtype=torch.float64)
    J[:, 0, 0] = fx / z
    J[:, 0, 2] = -fx * t[:, 0] / (z * z)
    J[:, 1, 1] = fy / z
    J[:, 1, 2] = -fy * t[:, 1] / (z * z)
    Wr = V[:3, :3].T
    T = J @ Wr
    cov2 = T @ Sigma @ T.transpose(1, 2) + 0.3 * torch.eye(2, dtype=torch.float64)
    conic = torch.linalg.inv(cov2)
    dirs = p["means3D"] - campos
    dirs = dirs / dirs.norm(dim=1, keepdim=True)
    rgb = sh_rgb(dirs, p["dc"], p["shs"])
    op = p["opacities"].reshape(P)

    keys, vals = orc.get("keys"), orc.get("vals").astype(np.int64)
    rng = orc.get("ranges").reshape(-1, 2)
    contrib = orc.get("contrib").reshape(H, W)
    img = torch.zeros(3, H, W, dtype=torch.float64)
    gx = orc.gx
    for tile in range(rng.shape[0]):
        s, e = int(rng[tile, 0]), int(rng[tile, 1])
        tx, ty = tile % gx, tile // gx
        ys, xs = np.mgrid[ty * 16:min(ty * 16 + 16, H), tx * 16:min(tx * 16 + 16, W)]
        ys, xs = ys.ravel(), xs.ravel()
        if xs.size == 0:
            continue
        ids = torch.from_numpy(vals[s:e])
        cnt = torch.from_numpy(contrib[ys, xs].astype(np.int64))
        if e > s:
            dx = px[ids][None, :] - torch.from_numpy(xs.astype(np.float64))[:, None]
            dy = py[ids][None, :] - torch.from_numpy(ys.astype(np.float64))[:, None]
            c = conic[ids]
            q = c[:, 0, 0] * dx * dx + 2 * c[:, 0, 1] * dx * dy + c[:, 1, 1] * dy * dy
            alpha = op[ids][None, :] * torch.exp(-0.5 * q)
            keep = (alpha.detach() >= 1 / 255) & (torch.arange(e - s)[None, :] < cnt[:, None])
            alpha = torch.where(keep, torch.clamp(alpha, max=0.99), torch.zeros_like(alpha))
            Tin = torch.cumprod(torch.cat([torch.ones(xs.size, 1, dtype=torch.float64),
                                           1 - alpha[:, :-1]], 1), 1)
            w = alpha * Tin
            col = w @ rgb[ids]  # (npix, 3)
            Tf = Tin[:, -1] * (1 - alpha[:, -1])
        else:
            col = torch.zeros(xs.size, 3, dtype=torch.float64)
            Tf = torch.ones(xs.size, dtype=torch.float64)
        bgt = torch.tensor(bg, dtype=torch.float64)
        val = col + Tf[:, None] * bgt[None, :]
        img[:, ys, xs] = val.T
    return img


@pytest.mark.parametrize("mode", [0, 1])
def test_adjoint_matches_torch_autograd(mode):
    scene, cam = scenes.fd_scene(40, seed=3)
    cam = scenes.look_at_camera((0.0, 0.0, -4.0), 48, 40)
    bg = (0.2, 0.1, 0.3)
    orc = Oracle(scene, cam, bucket=16, mode=mode, dtype=np.float64, bg=bg).forward()
    assert (orc.get("radii") > 0).sum() >= 20
    p = {k: torch.from_numpy(v.astype(np.float64)).requires_grad_(True) for k, v in scene.items()}
    img = render_torch(p, cam, orc, bg)
    np.testing.assert_allclose(img.detach().numpy(), orc.image(), atol=1e-9)
    w = np.random.default_rng(2).standard_normal(img.shape)
    (img * torch.from_numpy(w)).sum().backward()
    orc.backward(w)
    g = orc.grads()
    for k in ("means3D", "scales", "rotations", "opacities", "dc", "shs"):
        ref = p[k].grad.numpy()
        got = g[k].reshape(ref.shape)
        scale = np.abs(ref).max()
        np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-6 * scale, err_msg=k)
