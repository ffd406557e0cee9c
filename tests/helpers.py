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
                the FP32 restatement is, or within `tol` of the FP64 oracle itself, or within
                DRR_ABS_FLOOR of the image maximum of it (the exact value to float precision:
                on faint rays grazing a sharp tissue/air edge both float paths place samples to
                within an ulp, and either may land nearer the exact value);
      unexplained  the rest.
    Returns the counts, the mask of unexplained pixels and the worst error among flips."""
    err = drr_rel_err(got, ref32)
    bad = err > tol
    flip = bad & (np.asarray(steps_dev) != np.asarray(steps_cpu))
    g = np.asarray(got, np.float64)
    r32 = np.asarray(ref32, np.float64)
    r64 = np.asarray(ref64, np.float64)
    nearer = np.abs(g - r64) <= np.abs(r32 - r64)
    near64 = (drr_rel_err(got, ref64) <= tol) | (np.abs(g - r64) <= DRR_ABS_FLOOR * max(1.0, float(np.abs(r64).max())))
    drift = bad & ~flip & (nearer | near64)
    unexplained = bad & ~flip & ~(nearer | near64)
    return {"pixels": int(err.size), "over_tol": int(bad.sum()), "flips": int(flip.sum()),
            "ref_drift": int(drift.sum()), "ref_drift_within_tol_only": int((drift & ~nearer).sum()),
            "unexplained": int(unexplained.sum()),
            "max_err": float(err.max()), "max_err_not_flip": float(err[~flip].max()) if (~flip).any() else 0.0,
            "unexplained_mask": unexplained}


def describe_pixels(mask, got, ref32, ref64, limit: int = 5) -> str:
    """The first few masked pixels as (index, device, FP32 restatement, FP64 oracle)."""
    idx = np.argwhere(np.asarray(mask))[:limit]
    g, r32, r64 = (np.asarray(a, np.float64) for a in (got, ref32, ref64))
    return "; ".join(f"{tuple(int(k) for k in i)}: dev {g[tuple(i)]:.9g} fp32 {r32[tuple(i)]:.9g} "
                     f"fp64 {r64[tuple(i)]:.9g}" for i in idx)
