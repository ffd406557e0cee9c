"""Fixed-image downsampling restatement (oracle.downsample_image: ITK 5.1's DiscreteGaussian +
cubic B-spline resampling, as xReg's DownsampleImage uses them) pinned against independent
implementations: scipy's exponentially scaled Bessel functions for the discrete Gaussian,
scipy.ndimage's correlate1d (edge replication) and map_coordinates (cubic B-spline, mirror
boundary) for the filtering and resampling, and closed forms (constants, linear ramps)."""
import numpy as np
import pytest
from scipy import ndimage, special


@pytest.mark.parametrize("sigma", [1.0, 2.0, 4.0, 0.7])
def test_gaussian_kernel_is_the_discrete_gaussian(oracle, sigma):
    t = sigma * sigma
    k = oracle.gaussian_kernel(t)
    r = len(k) // 2
    n = np.abs(np.arange(-r, r + 1))
    ref = special.ive(n, t)                 # e^-t I_n(t)
    ref /= ref.sum()
    np.testing.assert_allclose(k, ref, rtol=2e-5, atol=1e-9)
    assert abs(k.sum() - 1.0) < 1e-12
    assert np.array_equal(k, k[::-1])
    # truncated where the untruncated mass first reaches 99%
    mass = special.ive(0, t) + 2 * np.cumsum(special.ive(np.arange(1, r + 1), t))
    assert mass[-1] >= 0.99 - 1e-6 and (r == 1 or mass[-2] < 0.99 + 1e-6)


def _scipy_downsample(img, factor, sigma, k=None):
    sm = img.astype(np.float64)
    if k is not None:
        sm = ndimage.correlate1d(sm, k, axis=1, mode="nearest")
        sm = ndimage.correlate1d(sm, k, axis=0, mode="nearest")
    sm = sm.astype(np.float32).astype(np.float64)
    orows, ocols = int(img.shape[0] * factor + 0.5), int(img.shape[1] * factor + 0.5)
    yy, xx = np.meshgrid(np.arange(orows) / factor, np.arange(ocols) / factor, indexing="ij")
    return ndimage.map_coordinates(sm, [yy, xx], order=3, mode="mirror", prefilter=True)


@pytest.mark.parametrize("factor", [0.5, 0.25, 0.3])
def test_downsample_matches_scipy_pipeline(oracle, factor):
    rng = np.random.default_rng(11)
    img = ndimage.gaussian_filter(rng.random((96, 80)), 2.0).astype(np.float32) + 0.1
    got = oracle.downsample_image(img, factor, spacing=(0.8, 0.8))
    k = oracle.gaussian_kernel((0.5 / factor) ** 2)
    ref = _scipy_downsample(img, factor, -1, k)
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-6 * float(np.abs(ref).max()))


def test_bspline_without_smoothing_matches_scipy(oracle):
    rng = np.random.default_rng(12)
    img = rng.random((40, 33)).astype(np.float32)
    got = oracle.downsample_image(img, 0.5, sigma=0.0)
    ref = _scipy_downsample(img, 0.5, 0.0)
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-6)
    # factor 1 without smoothing: the interpolating spline returns the pixels
    same = oracle.downsample_image(img, 1.0, sigma=0.0)
    np.testing.assert_allclose(same, img, rtol=0, atol=1e-6)


def test_constant_and_ramp(oracle):
    img = np.full((64, 64), 3.25, np.float32)
    out = oracle.downsample_image(img, 0.25)
    np.testing.assert_allclose(out, 3.25, rtol=1e-6)
    r, c = np.mgrid[0:128, 0:128]
    ramp = (0.5 + 0.01 * c + 0.02 * r).astype(np.float32)
    out = oracle.downsample_image(ramp, 0.5)
    i, j = np.mgrid[0:64, 0:64]
    want = 0.5 + 0.01 * (2 * j) + 0.02 * (2 * i)
    inner = (slice(16, 48), slice(16, 48))  # away from the edge replication and the mirror
    np.testing.assert_allclose(out[inner], want[inner], rtol=2e-6)


def test_planted_mistake_is_caught(oracle):
    """A resampling grid shifted by half a pixel (a plausible off-by-half) fails the scipy pin."""
    rng = np.random.default_rng(13)
    img = ndimage.gaussian_filter(rng.random((48, 48)), 1.5).astype(np.float32)
    got = oracle.downsample_image(img, 0.5, sigma=0.0)
    sm = img.astype(np.float64)
    yy, xx = np.meshgrid(np.arange(24) * 2 + 0.5, np.arange(24) * 2 + 0.5, indexing="ij")
    shifted = ndimage.map_coordinates(sm, [yy, xx], order=3, mode="mirror")
    assert np.abs(got - shifted).max() > 1e-3
