"""Seeded synthetic scenes for the textured-2DGS workloads.

Every scene is generated on the CPU with ``numpy.random.default_rng(seed)`` in
float32, so any consumer (GPU kernels, the CPU oracle) sees identical bits.

Workloads (sizes are the defaults; tests shrink them):
  1  one 256x256 view, 1,000 Gaussians, z~U[2,6], 4x4 RGB texels each,
     fx=fy=256, colours U[0,1]
  2  one 1920x1080 view, 500k Gaussians, z~U[2,20], SH degree 3 (48 coeffs
     ~N(0,0.1)), a 2M-texel budget split by projected area into 1x1..16x16
     grids, fx=fy=1080
  3  one 1600x1200 view, 1M untextured Gaussians (one texel each), fx=fy=1200
  4  1M Gaussians, SH degree 3, a 4M-texel budget; rendered from 16 cameras on
     a sphere of radius 8 looking at the origin (``make_orbit_views``), each
     at 512x512, 960x540, 1280x720 or 1920x1080 with fx=fy=H
  5  3M Gaussians, SH degree 3, an 8M-texel budget; rendered from 8 cameras at
     1920x1080 jittered around the identity pose (``make_views``)
Shared distributions: means xy~U[-1,1]; log-scales of the two in-plane axes
~U[-4,-2]; uniformly random unit quaternions; opacity = sigmoid(N(0,1));
uv0 = (0.5, 0.5); umap = ax1 / (6 s_x), vmap = ax2 / (6 s_y) so the +-3 sigma
footprint spans the chart; identity camera pose.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

SH_C0 = 0.28209479177387814


@dataclass
class SceneSpec:
    num_points: int
    img_w: int
    img_h: int
    focal: float
    z_range: tuple = (2.0, 6.0)
    sh_degree: int = -1          # -1: colours given directly
    texel_side: int = 4          # fixed grid side when texel_budget == 0
    texel_budget: int = 0        # >0: area-proportional grids of side 1..16
    channels: int = 3
    block_width: int = 16
    settings: int = 1 << 8
    background: tuple = (0.0, 0.0, 0.0)
    uv_scale: float = 6.0        # umap = ax1 / (uv_scale * s_x)


WORKLOADS: Dict[int, SceneSpec] = {
    1: SceneSpec(num_points=1000, img_w=256, img_h=256, focal=256.0, z_range=(2.0, 6.0),
                 texel_side=4),
    2: SceneSpec(num_points=500_000, img_w=1920, img_h=1080, focal=1080.0, z_range=(2.0, 20.0),
                 sh_degree=3, texel_budget=2_000_000),
    3: SceneSpec(num_points=1_000_000, img_w=1600, img_h=1200, focal=1200.0,
                 z_range=(2.0, 20.0), texel_side=1),
    4: SceneSpec(num_points=1_000_000, img_w=1920, img_h=1080, focal=1080.0,
                 z_range=(2.0, 20.0), sh_degree=3, texel_budget=4_000_000),
    5: SceneSpec(num_points=3_000_000, img_w=1920, img_h=1080, focal=1080.0,
                 z_range=(2.0, 20.0), sh_degree=3, texel_budget=8_000_000),
}

# (W, H) the orbit cameras of the multi-view workload draw from
ORBIT_RESOLUTIONS = [(512, 512), (960, 540), (1280, 720), (1920, 1080)]



def quat_axes_np(q: np.ndarray):
    """Rotation axes (columns of R) of (w, x, y, z) quaternions, float64."""
    q = q.astype(np.float64)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    a1 = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)], -1)
    a2 = np.stack([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)], -1)
    a3 = np.stack([2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)], -1)
    return a1, a2, a3


def _projected_area(means, scales, a1, a2, focal):
    """Area of the +-3 sigma corner AABB in pixels (identity camera), float64."""
    m = means.astype(np.float64)
    corners = []
    for s1 in (1.0, -1.0):
        for s2 in (1.0, -1.0):
            c = m + 3 * s1 * scales[:, :1] * a1 + 3 * s2 * scales[:, 1:2] * a2
            z = np.maximum(c[:, 2], 0.01)
            corners.append(np.stack([c[:, 0] / z * focal, c[:, 1] / z * focal], -1))
    cs = np.stack(corners, 1)
    ext = cs.max(1) - cs.min(1)
    return np.maximum(ext[:, 0] * ext[:, 1], 1e-6)


@dataclass
class Scene:
    """All host arrays of one view (float32 / int32, C-contiguous)."""

    means: np.ndarray
    scales: np.ndarray
    quats: np.ndarray
    opacities: np.ndarray
    colors: np.ndarray
    uv0: np.ndarray
    umap: np.ndarray
    vmap: np.ndarray
    texture: np.ndarray
    texture_dims: np.ndarray
    viewmat: np.ndarray
    c2w: np.ndarray
    background: np.ndarray
    target: np.ndarray
    fx: float
    fy: float
    cx: float
    cy: float
    img_h: int
    img_w: int
    block_width: int
    settings: int
    sh_coeffs: Optional[np.ndarray] = None
    sh_degree: int = -1
    extra: dict = field(default_factory=dict)

    @property
    def num_points(self) -> int:
        return int(self.means.shape[0])

    @property
    def texture_info(self):
        return (self.num_points, 1, int(self.texture.shape[1]))

    def viewdirs(self) -> np.ndarray:
        """Mean minus camera centre (the SH view direction), float32."""
        return (self.means - self.c2w[:3, 3][None, :]).astype(np.float32)

    def to_torch(self, device):
        import torch

        out = {}
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray):
                out[k] = torch.from_numpy(np.ascontiguousarray(v)).to(device)
        return out


def make_scene(workload: int = 1, seed: Optional[int] = None, num_points: Optional[int] = None,
               img_w: Optional[int] = None, img_h: Optional[int] = None,
               spec: Optional[SceneSpec] = None) -> Scene:
    """Generate one seeded scene (seed defaults to the workload number)."""
    sp = spec if spec is not None else WORKLOADS[workload]
    n = int(num_points if num_points is not None else sp.num_points)
    W = int(img_w if img_w is not None else sp.img_w)
    H = int(img_h if img_h is not None else sp.img_h)
    focal = sp.focal * (H / sp.img_h) if img_h is not None else sp.focal
    rng = np.random.default_rng(workload if seed is None else seed)
    f32 = np.float32

    means = np.empty((n, 3), f32)
    means[:, :2] = rng.uniform(-1.0, 1.0, (n, 2))
    means[:, 2] = rng.uniform(sp.z_range[0], sp.z_range[1], n)
    log_s = rng.uniform(-4.0, -2.0, (n, 2))
    scales = np.empty((n, 3), f32)
    scales[:, :2] = np.exp(log_s)
    scales[:, 2] = 1e-5 * scales[:, :2].mean(1)
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    quats = q.astype(f32)
    opac = (1.0 / (1.0 + np.exp(-rng.standard_normal((n, 1))))).astype(f32)

    a1, a2, _ = quat_axes_np(quats)
    sx = scales[:, :1].astype(np.float64)
    sy = scales[:, 1:2].astype(np.float64)
    umap = (a1 / (sp.uv_scale * sx)).astype(f32).reshape(n, 1, 3)
    vmap = (a2 / (sp.uv_scale * sy)).astype(f32).reshape(n, 1, 3)
    uv0 = np.full((n, 1, 2), 0.5, f32)

    if sp.texel_budget > 0:
        area = _projected_area(means, scales[:, :2].astype(np.float64), a1, a2, focal)
        side = np.clip(np.rint(np.sqrt(sp.texel_budget * area / area.sum())), 1, 16).astype(np.int64)
    else:
        side = np.full(n, sp.texel_side, np.int64)
    hw = side * side
    start = np.concatenate([[0], np.cumsum(hw)[:-1]])
    texture_dims = np.stack([side, side, start], 1).astype(np.int32)
    X = int(hw.sum())
    texture = rng.uniform(0.0, 1.0, (X, sp.channels)).astype(f32)

    sh_coeffs = None
    if sp.sh_degree >= 0:
        nb = (sp.sh_degree + 1) ** 2
        sh_coeffs = (0.1 * rng.standard_normal((n, nb, 3))).astype(f32)
        colors = np.zeros((n, 3), f32)  # filled by the SH evaluation at render time
    else:
        colors = rng.uniform(0.0, 1.0, (n, 3)).astype(f32)

    eye = np.eye(4, dtype=f32)
    target = rng.uniform(0.0, 1.0, (H, W, 3)).astype(f32)
    return Scene(means=means, scales=scales, quats=quats, opacities=opac, colors=colors,
                 uv0=uv0, umap=umap, vmap=vmap, texture=texture, texture_dims=texture_dims,
                 viewmat=eye.copy(), c2w=eye.copy(), background=np.asarray(sp.background, f32),
                 target=target, fx=float(focal), fy=float(focal), cx=W / 2.0, cy=H / 2.0,
                 img_h=H, img_w=W, block_width=sp.block_width, settings=sp.settings,
                 sh_coeffs=sh_coeffs, sh_degree=sp.sh_degree)


@dataclass
class View:
    """One camera of a multi-view batch (row-major 4x4 matrices, float32)."""
    viewmat: np.ndarray
    c2w: np.ndarray
    fx: float
    fy: float
    cx: float
    cy: float
    img_h: int
    img_w: int
    target: np.ndarray


def make_views(scene: Scene, num_views: int, seed: int = 7, max_angle_deg: float = 6.0,
               max_shift: float = 0.15,
               resolutions: Optional[list] = None) -> list:
    """Seeded cameras around the scene's camera: a random rotation of up to
    ``max_angle_deg`` about a random axis and a shift of up to ``max_shift`` in
    each coordinate, optionally with per-view resolutions (W, H) drawn from
    ``resolutions`` (focal scaled with the height).  Targets are U[0,1]."""
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(num_views):
        axis = rng.standard_normal(3)
        axis /= np.linalg.norm(axis)
        ang = np.deg2rad(rng.uniform(-max_angle_deg, max_angle_deg))
        K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
        R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * (K @ K)  # camera-to-world rotation
        t = rng.uniform(-max_shift, max_shift, 3)
        c2w = np.eye(4)
        c2w[:3, :3] = R
        c2w[:3, 3] = t
        viewmat = np.linalg.inv(c2w)
        W, H = scene.img_w, scene.img_h
        if resolutions:
            W, H = resolutions[int(rng.integers(0, len(resolutions)))]
        f = scene.fx * H / scene.img_h
        views.append(View(viewmat=viewmat.astype(np.float32), c2w=c2w.astype(np.float32),
                          fx=float(f), fy=float(f), cx=W / 2.0, cy=H / 2.0, img_h=int(H),
                          img_w=int(W),
                          target=rng.uniform(0.0, 1.0, (H, W, 3)).astype(np.float32)))
    return views


def _look_at(eye: np.ndarray, target: np.ndarray) -> np.ndarray:
    """Camera-to-world 4x4 (float64) at ``eye`` whose +z axis points at
    ``target``; x and y complete a right-handed frame (det R = +1)."""
    z = target - eye
    z = z / np.linalg.norm(z)
    ref = np.array([0.0, 1.0, 0.0]) if abs(z[1]) < 0.95 else np.array([1.0, 0.0, 0.0])
    x = np.cross(z, ref)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    c2w = np.eye(4)
    c2w[:3, 0], c2w[:3, 1], c2w[:3, 2], c2w[:3, 3] = x, y, z, eye
    return c2w


def make_orbit_views(num_views: int = 16, radius: float = 8.0, seed: int = 4,
                     resolutions: Optional[list] = None, jitter: float = 0.05) -> list:
    """Seeded cameras on a sphere of ``radius`` looking at the origin.

    Directions follow a Fibonacci lattice, each perturbed by N(0, ``jitter``)
    per coordinate and renormalised; the radius is scaled by U[0.95, 1.05] and
    the look-at point drawn from U[-0.1, 0.1]^3.  Each view draws its (W, H)
    from ``resolutions`` (default 512x512, 960x540, 1280x720, 1920x1080) and
    uses fx = fy = H, cx = W/2, cy = H/2.  Targets are U[0,1]."""
    rng = np.random.default_rng(seed)
    res = ORBIT_RESOLUTIONS if resolutions is None else resolutions
    golden = np.pi * (3.0 - np.sqrt(5.0))
    views = []
    for i in range(num_views):
        yy = 1.0 - 2.0 * (i + 0.5) / num_views
        r = np.sqrt(max(1.0 - yy * yy, 0.0))
        d = np.array([r * np.cos(golden * i), yy, r * np.sin(golden * i)])
        d = d + rng.normal(0.0, jitter, 3)
        d /= np.linalg.norm(d)
        eye = d * radius * rng.uniform(0.95, 1.05)
        c2w = _look_at(eye, rng.uniform(-0.1, 0.1, 3))
        W, H = res[int(rng.integers(0, len(res)))]
        views.append(View(viewmat=np.linalg.inv(c2w).astype(np.float32),
                          c2w=c2w.astype(np.float32), fx=float(H), fy=float(H), cx=W / 2.0,
                          cy=H / 2.0, img_h=int(H), img_w=int(W),
                          target=rng.uniform(0.0, 1.0, (H, W, 3)).astype(np.float32)))
    return views


def workload_views(workload: int, scene: Scene) -> list:
    """The cameras of a multi-view workload: 16 orbit views for 4, 8 views
    around the identity pose at the scene's 1920x1080 for 5."""
    if workload == 4:
        return make_orbit_views(16, radius=8.0, seed=4)
    if workload == 5:
        return make_views(scene, 8, seed=5)
    raise ValueError(f"workload {workload} has a single view")


def view_scene(scene: Scene, view: View) -> Scene:
    """The scene seen from ``view`` (shares every parameter array)."""
    import dataclasses

    return dataclasses.replace(scene, viewmat=view.viewmat, c2w=view.c2w, fx=view.fx,
                               fy=view.fy, cx=view.cx, cy=view.cy, img_h=view.img_h,
                               img_w=view.img_w, target=view.target)
