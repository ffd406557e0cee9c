"""Seeded synthetic workloads: ellipsoid phantoms, cameras and pose sets.

Phantom recipe (attenuation in 1/mm, air = 0): an outer soft-tissue ellipsoid (0.02, semi-axes
0.45 of the extent), a cortical shell (0.05, outer semi-axes 0.30 of the extent, two voxels thick)
and a trabecular interior (0.03 * (1 + U(-0.1, 0.1))). Volumes are returned z, y, x (x fastest),
which is ITK's buffer order.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import geometry as G

PHANTOM_SEED, POSE_SEED, NOISE_SEED = 1, 2, 3


def phantom(size_xyz, seed: int = PHANTOM_SEED) -> np.ndarray:
    X, Y, Z = (int(s) for s in size_xyz)
    rng = np.random.default_rng(seed)
    z = np.arange(Z, dtype=np.float32)
    y = np.arange(Y, dtype=np.float32)[:, None]
    x = np.arange(X, dtype=np.float32)[None, :]
    cx, cy, cz = (X - 1) / 2.0, (Y - 1) / 2.0, (Z - 1) / 2.0

    def ell(fx, fy, fz, shrink=0.0):
        # float32 terms of ((x-cx)/ax)^2 + ((y-cy)/ay)^2 + ((z-cz)/az)^2, summed in that order
        ax, ay, az = fx * X - shrink, fy * Y - shrink, fz * Z - shrink
        return ((x - cx) / ax) ** 2 + ((y - cy) / ay) ** 2, ((z - cz) / az) ** 2

    outer, shell_outer, inner = ell(0.45, 0.45, 0.45), ell(0.30, 0.30, 0.30), ell(0.30, 0.30, 0.30, shrink=2.0)
    vol = np.zeros((Z, Y, X), dtype=np.float32)
    # slice by slice (the noise stream is drawn in the same order as one (Z, Y, X) draw)
    for k in range(Z):
        sl = vol[k]
        sl[outer[0] + outer[1][k] <= 1.0] = 0.02
        sl[shell_outer[0] + shell_outer[1][k] <= 1.0] = 0.05
        noise = rng.uniform(-0.1, 0.1, size=(Y, X)).astype(np.float32)
        m = inner[0] + inner[1][k] <= 1.0
        sl[m] = (0.03 * (1.0 + noise))[m]
    return vol


@dataclass
class Workload:
    """One single-volume registration workload."""
    name: str
    vol: np.ndarray                  # z, y, x float32
    spacing: float
    cams: list                       # xrh Camera per view
    step: float
    gt: np.ndarray                   # 4x4 float64 cam->phys ground truth
    poses: np.ndarray                # [n, 4, 4] float64
    center: np.ndarray
    metric: int = 3
    smooth_ksize: int = 5
    patch_radius: int = 5
    extra: dict = field(default_factory=dict)

    @property
    def size_xyz(self):
        return (self.vol.shape[2], self.vol.shape[1], self.vol.shape[0])

    @property
    def idx_to_phys(self) -> np.ndarray:
        return G.idx_to_phys(self.spacing)


def random_poses(gt: np.ndarray, center, n: int, rot_deg: float, trans_mm: float, seed: int = POSE_SEED,
                 normal: bool = False) -> np.ndarray:
    rng = np.random.default_rng(seed)
    out = np.empty((n, 4, 4), dtype=np.float64)
    for i in range(n):
        if normal:
            r = np.deg2rad(rng.normal(0.0, rot_deg, size=3))
            t = rng.normal(0.0, trans_mm, size=3)
        else:
            r = np.deg2rad(rng.uniform(-rot_deg, rot_deg, size=3))
            t = rng.uniform(-trans_mm, trans_mm, size=3)
        out[i] = G.perturbed_pose(gt, center, r, t)
    return out


def multiview_extrinsics(focal: float, n_views: int, sep_deg: float):
    """float32 extrinsics (shared camera-world frame -> view camera frame) of views separated by
    sep_deg about the camera y axis through the isocentre (the volume's superior-inferior axis
    under the AP orientation). View 0 is the identity."""
    iso = focal / 2.0
    out = []
    for v in range(n_views):
        R = G.rot_y(np.deg2rad(sep_deg * v))
        E = G.rigid(None, (0, 0, -iso)) @ G.rigid(R.T) @ G.rigid(None, (0, 0, iso))
        out.append(E.astype(np.float32))
    return out


def multiview_cams(focal: float, rows: int, cols: int, px: float, n_views: int, sep_deg: float):
    """View 0 via CameraModel::setup(f, ...); views v > 0 via setup(intrins, extrins, ...) with
    the naive intrinsics and multiview_extrinsics, as an xReg multi-view set-up builds them."""
    cams = []
    K = G.naive_intrins(focal, rows, cols, px, px)
    for v, E in enumerate(multiview_extrinsics(focal, n_views, sep_deg)):
        cams.append(G.camera(focal, rows, cols, px, px) if v == 0 else G.camera_ke(K, E, rows, cols, px, px))
    return cams


def config(name: str, n_poses: int | None = None, small: bool = False) -> Workload:
    """Synthetic workloads by name: 'c1' (128x128x96 @2 mm, 192^2, single poses, NCC), 'c2'
    (256x256x200 @1 mm, 384^2, CMA-ES batches, patch gradient-NCC), 'c3' (the c2 volume, three
    views 30 degrees apart), 'c4' (512x512x400 @0.5 mm, 768^2, storage-precision study).
    small=True shrinks the volume and detector (same geometry) for fast parity tests."""
    focal = 1020.0
    if name == "c1":
        size, sp, det, px, step, n, metric = (128, 128, 96), 2.0, 192, 1.552, 1.0, 100, 0
    elif name == "c2":
        size, sp, det, px, step, n, metric = (256, 256, 200), 1.0, 384, 0.776, 0.5, 100, 3
    elif name == "c3":
        size, sp, det, px, step, n, metric = (256, 256, 200), 1.0, 384, 0.776, 0.5, 50, 1
    elif name == "c4":
        size, sp, det, px, step, n, metric = (512, 512, 400), 0.5, 768, 0.388, 0.25, 200, 0
    else:
        raise ValueError(name)
    if small:
        size = tuple(max(8, s // 4) for s in size)
        sp = sp * 4
        det = det // 4
        px = px * 4
        step = step * 4
    if n_poses is not None:
        n = n_poses
    vol = phantom(size)
    center = G.vol_center_phys(size, sp)
    gt = G.gt_pose(center, focal / 2.0)
    if name == "c3":
        cams = multiview_cams(focal, det, det, px, 3, 30.0)
    else:
        cams = [G.camera(focal, det, det, px, px)]
    if name == "c2":
        poses = random_poses(gt, center, n, 10.0, 15.0, normal=True)
    else:
        poses = random_poses(gt, center, n, 15.0, 20.0)
    return Workload(name=name + ("-small" if small else ""), vol=vol, spacing=sp, cams=cams, step=step, gt=gt,
                    poses=poses, center=center, metric=metric)


def noisy_fixed(drr: np.ndarray, seed: int = NOISE_SEED, frac: float = 0.01) -> np.ndarray:
    """Fixed image = DRR + N(0, frac * max)."""
    rng = np.random.default_rng(seed)
    return (drr + rng.normal(0.0, frac * float(drr.max()), size=drr.shape)).astype(np.float32)
