"""Host-side geometry in xReg's float conventions.

CameraModel::setup (lib/transforms/xregPerspectiveXform.cpp:200-254) and the inverse
extrinsics SE3Inv (lib/transforms/xregRigidUtils.cpp) restated in float32 so the values handed
to the device are the ones xReg itself would hold. Poses are composed in float64.
"""
from __future__ import annotations

import math

import numpy as np

from ._xrh import Camera

F = np.float32

FRAME_POS_Z, FRAME_NEG_Z, FRAME_ON_DETECTOR = 0, 1, 2


def _mat3_inverse_f32(m: np.ndarray) -> np.ndarray:
    """3x3 inverse in Eigen's cofactor order (every entry cofactor * (1/det)), float32."""
    m = m.astype(F)

    def cof(i, j):
        i1, i2, j1, j2 = (i + 1) % 3, (i + 2) % 3, (j + 1) % 3, (j + 2) % 3
        return F(m[i1, j1] * m[i2, j2]) - F(m[i1, j2] * m[i2, j1])

    c00, c10, c20 = cof(0, 0), cof(1, 0), cof(2, 0)
    det = F(F(F(c00 * m[0, 0]) + F(c10 * m[1, 0])) + F(c20 * m[2, 0]))
    invdet = F(F(1.0) / det)
    r = np.empty((3, 3), dtype=F)
    r[0] = [F(c00 * invdet), F(c10 * invdet), F(c20 * invdet)]
    r[1] = [F(cof(0, 1) * invdet), F(cof(1, 1) * invdet), F(cof(2, 1) * invdet)]
    r[2] = [F(cof(0, 2) * invdet), F(cof(1, 2) * invdet), F(cof(2, 2) * invdet)]
    return r


def naive_intrins(focal: float, rows: int, cols: int, row_sp: float, col_sp: float) -> np.ndarray:
    """MakeNaiveIntrins (xregPerspectiveXform.cpp:200-228) for the NEG_Z frame, float32."""
    K = np.eye(3, dtype=F)
    K[0, 0] = F(F(focal) / F(col_sp))
    K[1, 1] = F(F(focal) / F(row_sp))
    K[0, 0] *= F(-1)
    K[1, 1] *= F(-1)
    K[0, 2] = F((cols - 1) * 0.5)
    K[1, 2] = F((rows - 1) * 0.5)
    return K


def se3_inv_f32(T: np.ndarray) -> np.ndarray:
    """SE3Inv: [R^T | -R^T t], float32 with the reference's product order."""
    T = T.astype(F)
    R = T[:3, :3].T.copy()
    t = T[:3, 3]
    out = np.eye(4, dtype=F)
    out[:3, :3] = R
    for i in range(3):
        s = F(F(F(R[i, 0] * t[0]) + F(R[i, 1] * t[1])) + F(R[i, 2] * t[2]))
        out[i, 3] = F(F(-1) * s)
    return out


def camera(focal: float, rows: int, cols: int, row_sp: float, col_sp: float,
           extrins: np.ndarray | None = None) -> Camera:
    """An xrh_camera exactly as CameraModel::setup would fill it (NEG_Z frame)."""
    c = Camera()
    c.num_det_rows = int(rows)
    c.num_det_cols = int(cols)
    c.coord_frame = FRAME_NEG_Z
    c.focal_len = float(F(focal))
    kinv = _mat3_inverse_f32(naive_intrins(focal, rows, cols, row_sp, col_sp))
    for i in range(9):
        c.intrins_inv[i] = float(kinv.reshape(-1)[i])
    if extrins is None:
        einv = np.eye(4, dtype=F)
    else:
        einv = se3_inv_f32(np.asarray(extrins))
    for i in range(16):
        c.extrins_inv[i] = float(einv.reshape(-1)[i])
    # NEG_Z / POS_Z: pinhole = extrins_inv translation (xregPerspectiveXform.cpp:285-288)
    for i in range(3):
        c.pinhole[i] = float(einv[i, 3])
    c.det_row_spacing = float(F(row_sp))
    c.det_col_spacing = float(F(col_sp))
    # the float intrinsics / extrinsics CameraModel keeps (used by downsample_camera)
    c.K = naive_intrins(focal, rows, cols, row_sp, col_sp)
    c.E = np.eye(4, dtype=F) if extrins is None else np.asarray(extrins, dtype=F)
    return c


def camera_ke(K: np.ndarray, E: np.ndarray, rows: int, cols: int, row_sp: float, col_sp: float,
              coord_frame: int = FRAME_NEG_Z) -> Camera:
    """CameraModel::setup(intrins, extrins, nr, nc, rs, cs) (xregPerspectiveXform.cpp:300-330):
    focal length from FocalLenFromIntrins (:186-190), inverse intrinsics, SE3Inv extrinsics;
    pinhole = extrins_inv translation with the origin at the focal point (POS_Z / NEG_Z), else
    extrins_inv * [0, 0, focal_len] (origin on the detector)."""
    if coord_frame not in (FRAME_POS_Z, FRAME_NEG_Z, FRAME_ON_DETECTOR):
        raise ValueError(f"unknown camera coordinate frame {coord_frame}")
    K = np.asarray(K, dtype=F)
    c = Camera()
    c.num_det_rows = int(rows)
    c.num_det_cols = int(cols)
    c.coord_frame = int(coord_frame)
    f = F(F(abs(F(K[0, 0] * F(col_sp))) + abs(F(K[1, 1] * F(row_sp)))) / F(2))
    c.focal_len = float(f)
    kinv = _mat3_inverse_f32(K)
    for i in range(9):
        c.intrins_inv[i] = float(kinv.reshape(-1)[i])
    einv = se3_inv_f32(np.asarray(E))
    for i in range(16):
        c.extrins_inv[i] = float(einv.reshape(-1)[i])
    for i in range(3):
        if coord_frame == FRAME_ON_DETECTOR:
            # affine * (0, 0, f): the linear part (only column 2 is non-zero) then the translation
            c.pinhole[i] = float(F(F(einv[i, 2] * f) + einv[i, 3]))
        else:
            c.pinhole[i] = float(einv[i, 3])
    c.det_row_spacing = float(F(row_sp))
    c.det_col_spacing = float(F(col_sp))
    c.K = K.copy()
    c.E = np.asarray(E, dtype=F).copy()
    return c


def camera_arrays(c: Camera):
    """(kinv 3x3, einv 4x4, pinhole 3) float32 views of a Camera."""
    kinv = np.array(c.intrins_inv[:], dtype=F).reshape(3, 3)
    einv = np.array(c.extrins_inv[:], dtype=F).reshape(4, 4)
    ph = np.array(c.pinhole[:], dtype=F)
    return kinv, einv, ph


def lround(x: float) -> int:
    """std::lround: nearest integer, halves away from zero."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def downsample_camera(c: Camera, ds: float, force_even_dims: bool = False) -> Camera:
    """DownsampleCameraModel (xregPerspectiveXform.cpp:654-688): the float intrinsics scaled by
    ds (entries (0,0), (1,1), (0,2), (1,2)), detector size lround(n * ds) (made even on request),
    pixel spacing / ds, then CameraModel::setup(intrins, extrins, ...). Cameras made by
    camera() / camera_ke() carry their float K and E; for others K is recovered from the inverse
    intrinsics and the extrinsics from the inverse extrinsics."""
    ds32 = F(ds)
    if hasattr(c, "K"):
        K = np.asarray(c.K, dtype=F).copy()
        E = np.asarray(c.E, dtype=F)
    else:
        kinv, einv, _ = camera_arrays(c)
        K = np.linalg.inv(kinv.astype(np.float64)).astype(F)
        E = np.linalg.inv(einv.astype(np.float64)).astype(F)
    for (r, k) in ((0, 0), (1, 1), (0, 2), (1, 2)):
        K[r, k] = F(K[r, k] * ds32)
    rows = lround(float(F(F(c.num_det_rows) * ds32)))
    cols = lround(float(F(F(c.num_det_cols) * ds32)))
    if force_even_dims:
        rows -= rows % 2
        cols -= cols % 2
    return camera_ke(K, E, rows, cols, float(F(F(c.det_row_spacing) / ds32)), float(F(F(c.det_col_spacing) / ds32)))


# ------------------------------------------------------------------------------------------
# poses (float64)

def rot_x(a: float) -> np.ndarray:
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]], dtype=np.float64)


def rot_y(a: float) -> np.ndarray:
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], dtype=np.float64)


def rot_z(a: float) -> np.ndarray:
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], dtype=np.float64)


def rigid(R: np.ndarray | None = None, t=(0.0, 0.0, 0.0)) -> np.ndarray:
    T = np.eye(4, dtype=np.float64)
    if R is not None:
        T[:3, :3] = R
    T[:3, 3] = t
    return T


# Camera axes in volume axes for an antero-posterior view: detector columns along +x,
# detector rows along +z (superior-inferior), rays (camera -z) along +y.
R_AP = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, -1.0], [0.0, 1.0, 0.0]])


def exp_se3(x: np.ndarray) -> np.ndarray:
    """ExpSE3 of [w; v] (lib/transforms/xregRigidUtils.cpp:40-85), float64."""
    w = np.asarray(x[:3], dtype=np.float64)
    v = np.asarray(x[3:6], dtype=np.float64)
    W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]], dtype=np.float64)
    th = float(np.linalg.norm(w))
    T = np.eye(4)
    if th > 1.0e-14:
        A = np.eye(3) + (math.sin(th) / th) * W + ((1 - math.cos(th)) / th ** 2) * (W @ W)
        V = np.eye(3) + ((1 - math.cos(th)) / th ** 2) * W + ((th - math.sin(th)) / th ** 3) * (W @ W)
        T[:3, :3] = A
        T[:3, 3] = V @ v
    else:
        T[:3, 3] = v
    return T


def exp_se3_batch(X: np.ndarray) -> np.ndarray:
    """exp_se3 over rows of an (n, 6) array -> (n, 4, 4) float64; same formulas, vectorised."""
    X = np.asarray(X, dtype=np.float64).reshape(-1, 6)
    n = X.shape[0]
    w, v = X[:, :3], X[:, 3:6]
    W = np.zeros((n, 3, 3))
    W[:, 0, 1], W[:, 0, 2] = -w[:, 2], w[:, 1]
    W[:, 1, 0], W[:, 1, 2] = w[:, 2], -w[:, 0]
    W[:, 2, 0], W[:, 2, 1] = -w[:, 1], w[:, 0]
    WW = W @ W
    th = np.linalg.norm(w, axis=1)
    big = th > 1.0e-14
    ts = np.where(big, th, 1.0)
    a = np.where(big, np.sin(ts) / ts, 0.0)[:, None, None]
    b = np.where(big, (1 - np.cos(ts)) / ts ** 2, 0.0)[:, None, None]
    c = np.where(big, (ts - np.sin(ts)) / ts ** 3, 0.0)[:, None, None]
    eye = np.eye(3)[None]
    A = np.where(big[:, None, None], eye + a * W + b * WW, eye)
    V = np.where(big[:, None, None], eye + b * W + c * WW, eye)
    T = np.zeros((n, 4, 4))
    T[:, :3, :3] = A
    T[:, :3, 3] = np.einsum("nij,nj->ni", V, v)
    T[:, 3, 3] = 1.0
    return T


def gt_pose(vol_center_phys, src_to_iso: float, view_rot: np.ndarray = R_AP) -> np.ndarray:
    """cam->phys pose putting the isocenter (camera (0,0,-src_to_iso)) at the volume centre."""
    return rigid(None, vol_center_phys) @ rigid(view_rot) @ rigid(None, (0.0, 0.0, src_to_iso))


def perturbed_pose(gt: np.ndarray, vol_center_phys, rot_rad, trans_mm) -> np.ndarray:
    """Rotation (x, y, z Euler, radians) about the volume centre and a translation (mm),
    applied on the volume side of the gt pose."""
    c = np.asarray(vol_center_phys, dtype=np.float64)
    R = rot_z(rot_rad[2]) @ rot_y(rot_rad[1]) @ rot_x(rot_rad[0])
    D = rigid(None, c) @ rigid(R, trans_mm) @ rigid(None, -c)
    return D @ gt


def idx_to_phys(spacing, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """ITKImagePhysicalPointTransformsAsEigen for an identity direction (float32 4x4)."""
    s = np.broadcast_to(np.asarray(spacing, dtype=np.float64), (3,))
    M = np.eye(4, dtype=F)
    for i in range(3):
        M[i, i] = F(s[i])
        M[i, 3] = F(origin[i])
    return M


def vol_center_phys(size_xyz, spacing, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Physical centre at continuous index size/2 (RayCaster::xform_img_center_to_itk_phys,
    lib/ray_cast/xregRayCastInterface.cpp:200-224)."""
    s = np.broadcast_to(np.asarray(spacing, dtype=np.float64), (3,))
    return np.asarray(origin, dtype=np.float64) + s * (np.asarray(size_xyz, dtype=np.float64) / 2.0)


def mean_projection_distance(pose_est: np.ndarray, pose_gt: np.ndarray, pts_phys, focal: float) -> float:
    """Mean distance (detector mm) between the projections of physical points under two
    cam->phys poses (NEG_Z pinhole at the origin, detector at z = -focal)."""
    d = []
    for p in np.asarray(pts_phys, dtype=np.float64).reshape(-1, 3):
        ph = np.append(p, 1.0)
        a = np.linalg.solve(pose_est, ph)
        b = np.linalg.solve(pose_gt, ph)
        d.append(np.linalg.norm(focal * a[:2] / -a[2] - focal * b[:2] / -b[2]))
    return float(np.mean(d))
