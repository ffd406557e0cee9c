"""Shared helpers for the parity tests (numpy only)."""
from __future__ import annotations

import numpy as np

# Per-pixel DRR tolerance: relative 1e-4 (the stated bar), with an absolute floor for pixels
# whose value is ~0 (rays that miss or graze the phantom) expressed relative to the image max.
DRR_REL_TOL = 1.0e-4
DRR_ABS_FLOOR = 1.0e-6
# Similarity tolerance: absolute 1e-5.
SIM_ABS_TOL = 1.0e-5


def drr_rel_err(got: np.ndarray, ref: np.ndarray) -> np.ndarray:
    """Per-pixel |got-ref| / max(|ref|, floor * max|ref|)."""
    ref = np.asarray(ref, dtype=np.float64)
    got = np.asarray(got, dtype=np.float64)
    scale = np.maximum(np.abs(ref), DRR_ABS_FLOOR * max(1.0, float(np.abs(ref).max())))
    return np.abs(got - ref) / scale


def poses_f32(poses) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(poses, dtype=np.float32).reshape(-1, 16))
