"""Seeded synthetic scenes and edits for the per-frame ARAP path.

Gaussians follow the reference's activated SoA layout (GaussianView.cpp:43-153
after activation): means (N,3), rotations (N,4) as (w,x,y,z) unit quaternions,
scales (N,3) exp-activated, opacities (N,) sigmoid-activated, SH (N,48) with
coefficient-major interleaving (coef*3 + channel).

Workloads:
  * ``bar``      — 100k Gaussians in [-2,2]x[-.25,.25]^2, 256 nodes, pinned
                    x<-1.7, active x>1.7 twisted 45 deg about +x, 8 GN iterations,
                    one 800x800 view at distance 6 (FoV 50 deg).
  * ``capsules`` — 1M Gaussians inside a union of 9 capsules in the unit sphere,
                    2048 nodes, one capsule pinned, another translated 0.3 along a
                    seeded unit vector, 10 GN iterations, 4 views at 800x800
                    (ring of radius 3, azimuth 0/90/180/270 deg, elevation 15 deg).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Gaussians:
    pos: np.ndarray
    rot: np.ndarray
    scale: np.ndarray
    opacity: np.ndarray
    sh: np.ndarray

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    def take(self, order: np.ndarray) -> "Gaussians":
        return Gaussians(*(np.ascontiguousarray(a[order]) for a in (self.pos, self.rot, self.scale, self.opacity,
                                                                      self.sh)))


@dataclass
class Camera:
    view: np.ndarray      # (16,) column-major world->camera (x right, y down, z forward)
    proj: np.ndarray      # (16,) column-major world->clip (projection @ view)
    campos: np.ndarray    # (3,)
    tanfovx: float
    tanfovy: float
    W: int
    H: int


@dataclass
class Scene:
    name: str
    gaussians: Gaussians
    n_nodes: int
    knn_k: int
    grid: int
    gn_iters: int
    cameras: list
    edit: dict = field(default_factory=dict)


def random_gaussians(rng: np.random.Generator, pos: np.ndarray, log_scale_mean=-4.5, log_scale_std=0.5,
                     flat=False) -> Gaussians:
    N = pos.shape[0]
    if flat:
        ls = rng.normal(-4.0, 0.5, size=(N, 3))
        ls[:, 2] = rng.normal(-9.0, 0.3, size=N)
    else:
        ls = rng.normal(log_scale_mean, log_scale_std, size=(N, 3))
    q = rng.normal(size=(N, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    op = 1.0 / (1.0 + np.exp(-rng.normal(size=N)))
    sh = rng.normal(0.0, 0.2, size=(N, 48))
    return Gaussians(pos.astype(np.float32), q.astype(np.float32), np.exp(ls).astype(np.float32),
                     op.astype(np.float32), sh.astype(np.float32))


def look_at_camera(eye, target, fov_deg=50.0, W=800, H=800, up=(0.0, 0.0, 1.0), znear=0.01, zfar=100.0) -> Camera:
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    x = np.cross(f, np.asarray(up, np.float64))
    if np.linalg.norm(x) < 1e-8:
        x = np.cross(f, np.array([0.0, 1.0, 0.0]))
    x /= np.linalg.norm(x)
    y = np.cross(f, x)
    R = np.stack([x, y, f])            # rows: camera axes in world coordinates
    t = -R @ eye
    V = np.eye(4)
    V[:3, :3] = R
    V[:3, 3] = t
    tan_half = math.tan(math.radians(fov_deg) * 0.5)
    tanfovy = tan_half
    tanfovx = tan_half * W / H
    top, right = tanfovy * znear, tanfovx * znear
    Pm = np.zeros((4, 4))
    Pm[0, 0] = znear / right
    Pm[1, 1] = znear / top
    Pm[2, 2] = zfar / (zfar - znear)
    Pm[2, 3] = -(zfar * znear) / (zfar - znear)
    Pm[3, 2] = 1.0
    full = Pm @ V
    return Camera(V.T.reshape(16).astype(np.float32), full.T.reshape(16).astype(np.float32),
                  eye.astype(np.float32), float(tanfovx), float(tanfovy), W, H)


def ortho_camera(eye, target, ortho_scale=2.0, W=800, H=800, up=(0.0, 0.0, 1.0), znear=0.01, zfar=100.0,
                 fov_deg=50.0) -> Camera:
    """Orthographic camera: GaussianView's setOrthoCam(ortho_scale, ortho_scale)
    (GaussianView.cpp:1018-1019, sibr::orthographic) in the rasteriser's
    convention (x right, y down, z forward): x_ndc = x_view / ortho_scale,
    y_ndc = y_view / ortho_scale, w = 1. tan_fov is passed along unused."""
    c = look_at_camera(eye, target, fov_deg, W, H, up, znear, zfar)
    V = c.view.reshape(4, 4).T.astype(np.float64)
    Pm = np.zeros((4, 4))
    Pm[0, 0] = 1.0 / ortho_scale
    Pm[1, 1] = 1.0 / ortho_scale
    Pm[2, 2] = 1.0 / (zfar - znear)
    Pm[2, 3] = -znear / (zfar - znear)
    Pm[3, 3] = 1.0
    full = Pm @ V
    return Camera(c.view, full.T.reshape(16).astype(np.float32), c.campos, c.tanfovx, c.tanfovy, W, H)


def ring_cameras(n, radius, elevation_deg, fov_deg=50.0, W=800, H=800):
    cams = []
    el = math.radians(elevation_deg)
    for i in range(n):
        az = 2.0 * math.pi * i / n
        eye = radius * np.array([math.cos(el) * math.cos(az), math.cos(el) * math.sin(az), math.sin(el)])
        cams.append(look_at_camera(eye, (0.0, 0.0, 0.0), fov_deg, W, H))
    return cams


def bar_scene(seed=0, n=100_000, n_nodes=256, W=800, H=800) -> Scene:
    rng = np.random.default_rng(seed)
    pos = rng.uniform([-2.0, -0.25, -0.25], [2.0, 0.25, 0.25], size=(n, 3))
    g = random_gaussians(rng, pos)
    cam = look_at_camera((0.0, -6.0, 0.0), (0.0, 0.0, 0.0), 50.0, W, H)
    return Scene("bar", g, n_nodes, 8, 64, 8, [cam],
                 edit=dict(kind="twist", pinned=("x<", -1.7), active=("x>", 1.7), axis=(1.0, 0.0, 0.0),
                           angle_deg=45.0))


def _capsules(rng):
    caps = []
    for _ in range(9):
        r = rng.uniform(0.05, 0.3)
        while True:
            a = rng.uniform(-1, 1, 3)
            b = rng.uniform(-1, 1, 3)
            if np.linalg.norm(a) <= 1.0 - r and np.linalg.norm(b) <= 1.0 - r:
                break
        caps.append((a, b, r))
    return caps


def _in_capsule(p, a, b, r):
    ab = b - a
    t = np.clip(((p - a) @ ab) / max(ab @ ab, 1e-12), 0.0, 1.0)
    d = p - (a + t[:, None] * ab)
    return np.einsum("ij,ij->i", d, d) <= r * r


def capsule_scene(seed=0, n=1_000_000, n_nodes=2048, W=800, H=800, n_views=4) -> Scene:
    rng = np.random.default_rng(seed)
    caps = _capsules(rng)
    vols = np.array([math.pi * r * r * np.linalg.norm(b - a) + 4.0 / 3.0 * math.pi * r ** 3 for a, b, r in caps])
    counts = rng.multinomial(n, vols / vols.sum())
    pts = []
    for (a, b, r), c in zip(caps, counts):
        lo = np.minimum(a, b) - r
        hi = np.maximum(a, b) + r
        got = np.zeros((0, 3))
        while got.shape[0] < c:
            cand = rng.uniform(lo, hi, size=(max(2 * (c - got.shape[0]), 64), 3))
            got = np.concatenate([got, cand[_in_capsule(cand, a, b, r)]])
        pts.append(got[:c])
    pos = np.concatenate(pts)
    pos = pos[rng.permutation(n)]
    g = random_gaussians(rng, pos)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    cams = ring_cameras(n_views, 3.0, 15.0, 50.0, W, H)
    return Scene("capsules", g, n_nodes, 8, 64, 10, cams,
                 edit=dict(kind="translate", capsules=caps, pinned_capsule=0, active_capsule=1,
                           direction=u, distance=0.3))


def flat_scene(seed=0, n=3_000_000, n_nodes=4096, W=1920, H=1080, n_views=8) -> Scene:
    """The large high-frequency workload: the capsule union with flat Gaussians
    (one scale axis exp(N(-9, 0.3)), two exp(N(-4, 0.5))), 4096 nodes, a 128^3
    grid, 12 GN iterations, 8 views at 1920x1080, and a combined edit of the
    active capsule: translation by 0.3, twist by 30 degrees about the capsule
    axis and scaling by 1.3 about its centre."""
    rng = np.random.default_rng(seed)
    caps = _capsules(rng)
    vols = np.array([math.pi * r * r * np.linalg.norm(b - a) + 4.0 / 3.0 * math.pi * r ** 3 for a, b, r in caps])
    counts = rng.multinomial(n, vols / vols.sum())
    pts = []
    for (a, b, r), c in zip(caps, counts):
        lo, hi = np.minimum(a, b) - r, np.maximum(a, b) + r
        got = np.zeros((0, 3))
        while got.shape[0] < c:
            cand = rng.uniform(lo, hi, size=(max(2 * (c - got.shape[0]), 64), 3))
            got = np.concatenate([got, cand[_in_capsule(cand, a, b, r)]])
        pts.append(got[:c])
    pos = np.concatenate(pts)[rng.permutation(n)]
    g = random_gaussians(rng, pos, flat=True)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    a1, b1, _ = caps[1]
    axis = (b1 - a1) / max(np.linalg.norm(b1 - a1), 1e-12)
    cams = ring_cameras(n_views, 3.0, 15.0, 50.0, W, H)
    return Scene("flat", g, n_nodes, 8, 128, 12, cams,
                 edit=dict(kind="bend_twist_scale", capsules=caps, pinned_capsule=0, active_capsule=1, direction=u,
                           distance=0.3, axis=axis, angle_deg=30.0, scale=1.3))


def blocks_for(scene: Scene, node_pos: np.ndarray):
    """Blocks (lists of node ids), types (0 pinned, 1 active) for the scene's edit."""
    e = scene.edit
    if e["kind"] == "twist":
        x = node_pos[:, 0]
        pinned = np.nonzero(x < e["pinned"][1])[0]
        active = np.nonzero(x > e["active"][1])[0]
    else:
        a, b, r = e["capsules"][e["pinned_capsule"]]
        pinned = np.nonzero(_in_capsule(node_pos.astype(np.float64), a, b, r))[0]
        a, b, r = e["capsules"][e["active_capsule"]]
        active = np.nonzero(_in_capsule(node_pos.astype(np.float64), a, b, r))[0]
        active = np.setdiff1d(active, pinned)
    blocks, types = [], []
    if pinned.size:
        blocks.append(pinned.astype(np.int32)); types.append(0)
    if active.size:
        blocks.append(active.astype(np.int32)); types.append(1)
    return blocks, np.array(types, np.int32)


def rotate_about_axis(p, center, axis, radian):
    """PointRotateByAxis (helper.cpp:1077-1100): Rodrigues rotation about the
    axis through `center`, in float."""
    p = np.asarray(p, np.float32)
    axis = np.asarray(axis, np.float32)
    axis = axis / np.float32(np.linalg.norm(axis))
    c, s = np.float32(math.cos(radian)), np.float32(math.sin(radian))
    x, y, z = axis
    R = np.array([[x * x * (1 - c) + c, x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                  [y * x * (1 - c) + z * s, y * y * (1 - c) + c, y * z * (1 - c) - x * s],
                  [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, z * z * (1 - c) + c]], np.float32)
    center = np.asarray(center, np.float32)
    return ((p - center) @ R.T + center).astype(np.float32)


def frame_aims(scene: Scene, node_pos: np.ndarray, blocks, types, fraction: float) -> np.ndarray:
    """Aims for one drag frame (UpdateAimPosition*, GaussianView.cpp:2920-2983):
    every aim starts at the current node position (ReloadAimPositions) and the
    active block moves by `fraction` of the whole edit."""
    aims = np.array(node_pos, np.float32, copy=True)
    e = scene.edit
    for blk, t in zip(blocks, types):
        if t != 1:
            continue
        if e["kind"] == "twist":
            center = node_pos[blk].astype(np.float32).mean(axis=0)
            aims[blk] = rotate_about_axis(node_pos[blk], center, e["axis"], math.radians(e["angle_deg"]) * fraction)
        elif e["kind"] == "bend_twist_scale":
            # UpdateAimPosition, then UpdateAimPositionTwist, then UpdateAimPositionScale
            p = node_pos[blk] + np.float32(e["distance"] * fraction) * e["direction"].astype(np.float32)
            center = p.astype(np.float32).mean(axis=0)
            p = rotate_about_axis(p, center, e["axis"], math.radians(e["angle_deg"]) * fraction)
            center = p.astype(np.float32).mean(axis=0)
            k = np.float32(1.0 + (e["scale"] - 1.0) * fraction)
            aims[blk] = (center + k * (p - center)).astype(np.float32)
        else:
            aims[blk] = node_pos[blk] + np.float32(e["distance"] * fraction) * e["direction"].astype(np.float32)
    return aims
