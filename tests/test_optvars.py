"""Pose parameterisations (SE3OptVars family) and the inter-frame sandwich, against closed
forms and independent rotation code (scipy)."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from xreg_amd import geometry as G
from xreg_amd import optvars as O
from xreg_amd import regi


def _rot(axis, a):
    T = np.eye(4)
    T[:3, :3] = Rotation.from_euler(axis, a).as_matrix()
    return T


def _tr(v):
    T = np.eye(4)
    T[:3, 3] = v
    return T


def test_euler_default_order():
    x = np.array([0.1, -0.2, 0.3, 4.0, -5.0, 6.0])
    want = _tr([0, 0, 6.0]) @ _tr([0, -5.0, 0]) @ _tr([4.0, 0, 0]) @ _rot("z", 0.3) @ _rot("y", -0.2) @ _rot("x", 0.1)
    np.testing.assert_allclose(O.SE3OptVarsEuler()(x), want, atol=1e-12)


def test_euler_custom_order_and_validation():
    x = np.array([0.1, -0.2, 0.3, 4.0, -5.0, 6.0])
    # translations first, then Rx, Ry, Rz
    v = O.SE3OptVarsEuler(rot_x_idx=3, rot_y_idx=4, rot_z_idx=5, trans_x_idx=0, trans_y_idx=1, trans_z_idx=2)
    want = _tr([4.0, 0, 0]) @ _tr([0, -5.0, 0]) @ _tr([0, 0, 6.0]) @ _rot("x", 0.1) @ _rot("y", -0.2) @ _rot("z", 0.3)
    np.testing.assert_allclose(v(x), want, atol=1e-12)
    with pytest.raises(ValueError):
        O.SE3OptVarsEuler(0, 0, 1, 2, 3, 4)


def test_so3_and_translation_only():
    w = np.array([0.2, -0.1, 0.4])
    np.testing.assert_allclose(O.SO3OptVarsLieAlg()(w)[:3, :3], Rotation.from_rotvec(w).as_matrix(), atol=1e-12)
    np.testing.assert_allclose(O.SO3OptVarsEuler()(w), _rot("z", 0.4) @ _rot("y", -0.1) @ _rot("x", 0.2), atol=1e-12)
    np.testing.assert_allclose(O.SO3OptVarsOnly(1)([0.3]), _rot("y", 0.3), atol=1e-12)
    t = O.SE3OptVarsTransOnly(True, False, True)
    assert t.num_params == 2
    np.testing.assert_allclose(t([1.5, -2.0]), _tr([1.5, 0, -2.0]))


def test_inter_frame_sandwich():
    rng = np.random.default_rng(3)
    guess = G.gt_pose(np.array([10.0, 20.0, 30.0]), 510.0)
    I = G.rigid(Rotation.from_rotvec(rng.normal(0, 0.3, 3)).as_matrix(), rng.normal(0, 20, 3))
    x = rng.normal(0, 0.1, 6)
    T = O.SE3OptVarsLieAlg()(x)
    a = O.InterFramePose(guess, None, I, inter_frame_wrt_vol=True)(x)
    b = O.InterFramePose(guess, None, I, inter_frame_wrt_vol=False)(x)
    np.testing.assert_allclose(a, I @ T @ np.linalg.inv(I) @ guess, atol=1e-9)
    np.testing.assert_allclose(b, guess @ I @ T @ np.linalg.inv(I), atol=1e-9)
    # x = 0 gives the guess in both conventions
    np.testing.assert_allclose(O.InterFramePose(guess, None, I, True)(np.zeros(6)), guess, atol=1e-9)
    np.testing.assert_allclose(O.InterFramePose(guess, None, I, False)(np.zeros(6)), guess, atol=1e-9)


def test_pose_param_is_the_rotation_centre_sandwich():
    c = np.array([64.0, 64.0, 50.0])
    init = G.gt_pose(c, 510.0)
    X = np.random.default_rng(4).normal(0, [0.05] * 3 + [5.0] * 3, (5, 6))
    P = regi.PoseParam(init, c)
    B = P(X)
    for x, b in zip(X, B):
        np.testing.assert_allclose(b, G.rigid(None, c) @ G.exp_se3(x) @ G.rigid(None, -c) @ init, atol=1e-9)
    assert P.num_params == 6
    P3 = regi.PoseParam(init, c, O.SE3OptVarsTransOnly())
    assert P3.num_params == 3
    np.testing.assert_allclose(P3(np.array([1.0, 2.0, 3.0])), _tr([1, 2, 3]) @ init, atol=1e-9)
