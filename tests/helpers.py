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


def attribute_drr_errors(got, ref32, ref64, steps_dev, steps_cpu, tol: float = DRR_REL_TOL) -> dict:
    """Splits the pixels where a DRR differs from the FP32 CPU restatement by more than `tol`
    (drr_rel_err) into
      flips     the two set-ups give the ray a different step count n (one sample more or less,
                ~1/n of the pixel: the discontinuity of floor(chord / step));
      ref_drift same step count, and the device value is at least as close to the FP64 oracle as
                the FP32 restatement is (the restatement's own float drift, e.g. its sample
                positions accumulated by repeated addition, is the larger error);
      unexplained  the rest.
    Returns the counts, the mask of unexplained pixels and the worst error among flips."""
    err = drr_rel_err(got, ref32)
    bad = err > tol
    flip = bad & (np.asarray(steps_dev) != np.asarray(steps_cpu))
    g = np.asarray(got, np.float64)
    r32 = np.asarray(ref32, np.float64)
    r64 = np.asarray(ref64, np.float64)
    closer = np.abs(g - r64) <= np.abs(r32 - r64)
    drift = bad & ~flip & closer
    unexplained = bad & ~flip & ~closer
    return {"pixels": int(err.size), "over_tol": int(bad.sum()), "flips": int(flip.sum()),
            "ref_drift": int(drift.sum()), "unexplained": int(unexplained.sum()),
            "max_err": float(err.max()), "max_err_not_flip": float(err[~flip].max()) if (~flip).any() else 0.0,
            "unexplained_mask": unexplained}
