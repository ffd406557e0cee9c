"""Pins the oracle's geometry pieces (CameraModel, RayRectIntersect, ITK trilinear, Eigen sum)
against closed forms, brute force and independent numpy arithmetic."""
import math

import numpy as np
import pytest


def test_camera_matches_product_setup_bitwise(oracle):
    from xreg_amd import geometry as G

    for (f, r, c, rs, cs) in [(1020.0, 384, 384, 0.776, 0.776), (1020.0, 192, 192, 1.552, 1.552),
                              (900.0, 97, 131, 0.5, 0.7)]:
        o = oracle.cam_naive(f, r, c, rs, cs)
        p = G.camera(f, r, c, rs, cs)
        assert np.array_equal(np.array(o.kinv[:], np.float32), np.array(p.intrins_inv[:], np.float32))
        assert np.array_equal(np.array(o.einv[:], np.float32), np.array(p.extrins_inv[:], np.float32))


@pytest.mark.parametrize("rows,cols,rs,cs", [(49, 49, 0.776, 0.776), (48, 64, 1.0, 0.5)])
def test_detector_points_closed_form(oracle, rows, cols, rs, cs):
    """ind_pt_to_phys_det_pt for the naive NEG_Z camera is ((c-cx) cs, (r-cy) rs, -f)."""
    f = 1020.0
    cam = oracle.cam_naive(f, rows, cols, rs, cs)
    cx, cy = (cols - 1) / 2.0, (rows - 1) / 2.0
    for (c, r) in [(0, 0), (cols - 1, rows - 1), (cx, cy), (3, rows - 2)]:
        p = oracle.cam_ind_to_det_f32(cam, c, r).astype(np.float64)
        want = np.array([(c - cx) * cs, (r - cy) * rs, -f])
        assert np.allclose(p, want, rtol=0, atol=2e-4 * f), (c, r, p, want)
    # the principal point maps onto the optical axis
    p = oracle.cam_ind_to_det_f32(cam, cx, cy)
    assert abs(p[0]) < 1e-4 and abs(p[1]) < 1e-4 and p[2] == np.float32(-f)


@pytest.mark.parametrize("frame", [0, 2])
def test_detector_points_other_frames(oracle, frame):
    """Positive-focal intrinsics (MakeNaiveIntrins without the NEG_Z flip): with the origin at
    the focal point (POS_Z) the detector is the plane z = +f and the pinhole the origin; with the
    origin on the detector the detector is z = 0 and the pinhole (0, 0, f)."""
    from xreg_amd import geometry as G

    f, rows, cols, rs, cs = 1020.0, 41, 37, 0.6, 0.8
    K = G.naive_intrins(f, rows, cols, rs, cs)
    K[0, 0] *= -1
    K[1, 1] *= -1
    cam = G.camera_ke(K, np.eye(4), rows, cols, rs, cs, coord_frame=frame)
    oc = oracle.cam_f(cam)
    cx, cy = (cols - 1) / 2.0, (rows - 1) / 2.0
    z = f if frame == 0 else 0.0
    for (c, r) in [(0, 0), (cols - 1, rows - 1), (cx, cy), (5, 30)]:
        p = oracle.cam_ind_to_det_f32(oc, c, r).astype(np.float64)
        want = np.array([(c - cx) * cs, (r - cy) * rs, z])
        assert np.allclose(p, want, rtol=0, atol=2e-4 * f), (c, r, p, want)
    assert np.allclose(np.array(cam.pinhole[:]), [0.0, 0.0, 0.0 if frame == 0 else f], atol=1e-3)


def test_detector_points_with_extrinsics(oracle):
    from xreg_amd import geometry as G
    from xreg_amd import synth

    cams = synth.multiview_cams(1020.0, 33, 33, 2.0, 3, 30.0)
    for v, cam in enumerate(cams):
        oc = oracle.cam_f(cam)
        p = oracle.cam_ind_to_det_f32(oc, 16.0, 16.0).astype(np.float64)
        # view v rotates the principal ray about the isocentre by 30 v degrees around y
        th = math.radians(30.0 * v)
        R = G.rot_y(th)
        want = np.array([0.0, 0.0, -510.0]) + R @ np.array([0.0, 0.0, -510.0])
        assert np.allclose(p, want, atol=5e-3), (v, p, want)
        # the pinhole sits on the other side of the isocentre
        ph = np.array(cam.pinhole[:], dtype=np.float64)
        assert np.allclose(ph, np.array([0, 0, -510.0]) - R @ np.array([0, 0, -510.0]), atol=5e-3)


def _brute_segment(mn, mx, p0, d, n=200001):
    t = np.linspace(0.0, 1.0, n)
    pts = p0[None, :] + t[:, None] * d[None, :]
    inside = np.all((pts >= mn - 1e-9) & (pts <= mx + 1e-9), axis=1)
    if not inside.any():
        return False, None, None
    return True, t[inside].min(), t[inside].max()


def test_ray_rect_against_brute_force(oracle):
    rng = np.random.default_rng(5)
    mn = np.zeros(3)
    mx = np.array([31.0, 20.0, 11.0])
    hits = 0
    for _ in range(200):
        p0 = rng.uniform(-40, 70, 3)
        q = rng.uniform(-10, 45, 3)
        d = q - p0
        hit, t0, t1 = oracle.ray_rect_f32(mn, mx, p0, d)
        bh, b0, b1 = _brute_segment(mn, mx, p0, d)
        if bh and (b1 - b0) > 1e-3:
            hits += 1
            assert hit
            assert abs(t0 - b0) < 2e-5 and abs(t1 - b1) < 2e-5
        elif not bh:
            assert not hit or (t1 - t0) < 1e-4
    assert hits > 20


def test_ray_rect_parallel_components(oracle):
    mn, mx = np.zeros(3), np.array([10.0, 10.0, 10.0])
    # parallel to the x slab, inside it -> hits
    hit, t0, t1 = oracle.ray_rect_f32(mn, mx, [5, -5, 5], [0, 20, 0])
    assert hit and abs(t0 - 0.25) < 1e-6 and abs(t1 - 0.75) < 1e-6
    # parallel to the x slab, outside it -> misses
    hit, _, _ = oracle.ray_rect_f32(mn, mx, [11, -5, 5], [0, 20, 0])
    assert not hit
    # components with |d| <= 1e-8 count as parallel
    hit, _, _ = oracle.ray_rect_f32(mn, mx, [-1, -5, 5], [5e-9, 20, 0])
    assert not hit
    # segment limit: the box lies beyond the segment end
    hit, _, _ = oracle.ray_rect_f32(mn, mx, [5, -30, 5], [0, 20, 0])
    assert not hit


def test_itk_interp_linear_volume_exact(oracle):
    """Trilinear interpolation reproduces a field linear in the index exactly; a transposed
    weight or a dropped term breaks it."""
    X, Y, Z = 9, 7, 5
    k, j, i = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    a, b, c, d = 0.5, -1.25, 2.0, 3.0
    vol = (a * i + b * j + c * k + d).astype(np.float32)
    rng = np.random.default_rng(1)
    for _ in range(300):
        p = rng.uniform([0, 0, 0], [X - 1, Y - 1, Z - 1]).astype(np.float32)
        want = a * float(p[0]) + b * float(p[1]) + c * float(p[2]) + d
        assert abs(oracle.itk_interp(vol, p) - want) < 1e-9 * 50


def test_itk_interp_voxel_centres_and_borders(oracle):
    rng = np.random.default_rng(2)
    vol = rng.uniform(0, 1, (4, 5, 6)).astype(np.float32)
    for _ in range(20):
        i, j, k = rng.integers(0, 6), rng.integers(0, 5), rng.integers(0, 4)
        assert oracle.itk_interp(vol, [i, j, k]) == float(vol[k, j, i])
    # below 0: clamped base, non-positive distance -> base voxel
    assert oracle.itk_interp(vol, [-0.3, 0, 0]) == float(vol[0, 0, 0])
    # past the last index: upper neighbour clamped to the last voxel
    assert oracle.itk_interp(vol, [5.4, 4, 3]) == float(vol[3, 4, 5])
    # mid-edge along x only
    want = 0.5 * (float(vol[1, 2, 3]) + float(vol[1, 2, 4]))
    assert abs(oracle.itk_interp(vol, [3.5, 2, 1]) - want) < 1e-12


def _eigen_sum_emulated(x):
    x = np.asarray(x, dtype=np.float32)
    n = x.size
    al2 = (n // 8) * 8
    al = (n // 4) * 4
    if al == 0:
        r = x[0]
        for v in x[1:]:
            r = np.float32(r + v)
        return float(r)
    p0 = x[0:4].copy()
    if al > 4:
        p1 = x[4:8].copy()
        for i in range(8, al2, 8):
            p0 = (p0 + x[i:i + 4]).astype(np.float32)
            p1 = (p1 + x[i + 4:i + 8]).astype(np.float32)
        p0 = (p0 + p1).astype(np.float32)
        if al > al2:
            p0 = (p0 + x[al2:al2 + 4]).astype(np.float32)
    r = np.float32(np.float32(p0[0] + p0[2]) + np.float32(p0[1] + p0[3]))
    for v in x[al:]:
        r = np.float32(r + v)
    return float(r)


@pytest.mark.parametrize("n", [1, 3, 4, 5, 7, 8, 9, 12, 15, 16, 17, 1001])
def test_eigen_sum_restatement(oracle, n):
    x = np.random.default_rng(n).normal(size=n).astype(np.float32)
    assert oracle.eigen_sum_f32(x) == _eigen_sum_emulated(x)
    assert abs(oracle.eigen_sum_f32(x) - math.fsum(x.astype(np.float64))) < 1e-5 * max(1, n)
    ints = np.arange(n, dtype=np.float32)
    assert oracle.eigen_sum_f32(ints) == n * (n - 1) / 2
