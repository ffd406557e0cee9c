"""Optimiser parameterisations of a rigid pose and the inter-frame sandwich.

SE3OptVars and its subclasses (lib/regi/xregSE3OptVars.h:44-300, .cpp:31-300) map an
optimiser vector to a 4x4 rigid transform; Intensity2D3DRegi::apply_inter_transforms_for_obj_fn
(lib/regi/interfaces_2d_3d/xregIntensity2D3DRegi.cpp:1027-1099) places that transform between
a registration guess and an intermediate frame:

* intermediate frame given w.r.t. the volume:  pose = I * T(x) * I^-1 * guess
* otherwise:                                     pose = guess * I * T(x) * I^-1

All arithmetic is float64 (the reference composes FrameTransforms in float; the result is
handed to the device as a float64 pose and rounded there).
"""
from __future__ import annotations

import math

import numpy as np

from . import geometry as G


def _rx(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1]], dtype=np.float64)


def _ry(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]], dtype=np.float64)


def _rz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0, 0], [s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], dtype=np.float64)


def _t(axis):
    def f(v):
        T = np.eye(4)
        T[axis, 3] = v
        return T
    return f


class SE3OptVars:
    num_params = 6

    def __call__(self, x: np.ndarray) -> np.ndarray:
        raise NotImplementedError

    def batch(self, X: np.ndarray) -> np.ndarray:
        X = np.asarray(X, dtype=np.float64).reshape(-1, self.num_params)
        return np.stack([self(x) for x in X])


class SE3OptVarsLieAlg(SE3OptVars):
    """ExpSE3 of [w; v] (xregSE3OptVars.cpp:127-137)."""
    num_params = 6

    def __call__(self, x):
        return G.exp_se3(np.asarray(x, dtype=np.float64))

    def batch(self, X):
        return G.exp_se3_batch(X)


class SE3OptVarsEuler(SE3OptVars):
    """Product of the six elementary transforms in a chosen order (xregSE3OptVars.cpp:31-80).
    Parameters are (rot x, rot y, rot z, trans x, trans y, trans z); *_idx is each factor's
    position in the product, left to right. The default gives T = Tz * Ty * Tx * Rz * Ry * Rx."""
    num_params = 6

    def __init__(self, rot_x_idx=5, rot_y_idx=4, rot_z_idx=3, trans_x_idx=2, trans_y_idx=1, trans_z_idx=0):
        idx = [rot_x_idx, rot_y_idx, rot_z_idx, trans_x_idx, trans_y_idx, trans_z_idx]
        if sorted(idx) != list(range(6)):
            raise ValueError("Euler factor positions must be a permutation of 0..5")
        fns = [_rx, _ry, _rz, _t(0), _t(1), _t(2)]
        self.order = [None] * 6  # (function, parameter index) per product position
        for p, pos in enumerate(idx):
            self.order[pos] = (fns[p], p)

    def __call__(self, x):
        T = np.eye(4)
        for fn, p in self.order:
            T = T @ fn(float(x[p]))
        return T


class SE3OptVarsTransOnly(SE3OptVars):
    """Translation along the selected axes (xregSE3OptVars.cpp:142-178)."""

    def __init__(self, use_x=True, use_y=True, use_z=True):
        self.axes = [a for a, u in enumerate((use_x, use_y, use_z)) if u]
        if not self.axes:
            raise ValueError("at least one translation axis")
        self.num_params = len(self.axes)

    def __call__(self, x):
        T = np.eye(4)
        for k, a in enumerate(self.axes):
            T[a, 3] = float(x[k])
        return T


class SO3OptVarsLieAlg(SE3OptVars):
    """ExpSO3 of w (xregSE3OptVars.cpp:203-213)."""
    num_params = 3

    def __call__(self, x):
        T = G.exp_se3(np.concatenate([np.asarray(x, dtype=np.float64), np.zeros(3)]))
        T[:3, 3] = 0.0
        return T


class SO3OptVarsEuler(SE3OptVars):
    """Three elementary rotations in a chosen order (xregSE3OptVars.cpp:215-262); default
    Rz * Ry * Rx."""
    num_params = 3

    def __init__(self, rot_x_idx=2, rot_y_idx=1, rot_z_idx=0):
        idx = [rot_x_idx, rot_y_idx, rot_z_idx]
        if sorted(idx) != [0, 1, 2]:
            raise ValueError("Euler factor positions must be a permutation of 0..2")
        fns = [_rx, _ry, _rz]
        self.order = [None] * 3
        for p, pos in enumerate(idx):
            self.order[pos] = (fns[p], p)

    def __call__(self, x):
        T = np.eye(4)
        for fn, p in self.order:
            T = T @ fn(float(x[p]))
        return T


class SO3OptVarsOnly(SE3OptVars):
    """One rotation about x, y or z (xregSE3OptVars.cpp:264-300)."""
    num_params = 1

    def __init__(self, axis: int):
        self.axis = int(axis)
        self.fn = (_rx, _ry, _rz)[axis]

    def __call__(self, x):
        return self.fn(float(np.asarray(x).reshape(-1)[0]))


class InterFramePose:
    """Optimiser vector -> cam->phys pose through the inter-frame sandwich."""

    def __init__(self, guess: np.ndarray, opt_vars: SE3OptVars | None = None,
                 inter_frame: np.ndarray | None = None, inter_frame_wrt_vol: bool = True):
        self.vars = opt_vars if opt_vars is not None else SE3OptVarsLieAlg()
        I = np.eye(4) if inter_frame is None else np.asarray(inter_frame, dtype=np.float64)
        g = np.asarray(guess, dtype=np.float64)
        self.guess, self.inter_frame, self.inter_frame_wrt_vol = g.copy(), I.copy(), bool(inter_frame_wrt_vol)
        Iinv = np.linalg.inv(I)
        if inter_frame_wrt_vol:
            self.pre, self.post = I, Iinv @ g
        else:
            self.pre, self.post = g @ I, Iinv
        self.num_params = self.vars.num_params

    def __call__(self, X: np.ndarray) -> np.ndarray:
        X = np.asarray(X, dtype=np.float64)
        if X.ndim == 1:
            return self.pre @ self.vars(X) @ self.post
        return self.pre @ self.vars.batch(X) @ self.post
