"""HU -> linear attenuation restatement (oracle.hu_to_lin_att) pinned by its closed form
max((hu - hu_lower) * (mu_water - mu_air) * 1e-3, 0) and its special values."""
import numpy as np


def test_hu_closed_form(oracle):
    rng = np.random.default_rng(3)
    hu = np.concatenate([rng.uniform(-1500, 3000, 10000), [-1000.0, -1000.5, 0.0, 1000.0, -3000.0]]).astype(np.float32)
    got = oracle.hu_to_lin_att(hu)
    s = (0.02683 - 0.02485e-4) * 1e-3
    want = np.maximum((hu.astype(np.float64) + 1000.0) * s, 0.0)
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-12)
    assert got[hu <= -1000.0].max() == 0.0  # at and below the lower bound: exactly zero
    assert abs(got[-3] - 1000.0 * s) < 1e-9  # water: 1000 HU above the lower bound


def test_hu_other_lower_bound_and_nan(oracle):
    hu = np.array([-130.0, -131.0, 0.0, np.nan], np.float32)
    got = oracle.hu_to_lin_att(hu, hu_lower=-130.0)
    s = (0.02683 - 0.02485e-4) * 1e-3
    assert got[0] == 0.0 and got[1] == 0.0
    assert abs(got[2] - 130.0 * s) < 1e-9
    assert np.isnan(got[3])  # std::max(NaN, 0.0) keeps the NaN
