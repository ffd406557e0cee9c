"""Pins the oracle's nearest-neighbour, windowed-sinc and cubic B-spline interpolators (the ITK
interpolators RayCasterLineIntCPU can use, lib/ray_cast/xregRayCastLineIntCPU.cpp:115-147)
against independent routines and closed forms:
  * nearest neighbour: numpy indexing at floor(x + 0.5) away from half-integers, and ITK's
    documented RoundHalfIntegerUp on exact half-integers (2.5 -> 3, not banker's 2);
  * windowed sinc: a numpy evaluation with np.sinc (sin(pi x) / (pi x)) for both the sinc and
    the Lanczos window, taps floor(x) - 3 .. floor(x) + 4, zeros outside the volume; the
    unit impulse at integer indices;
  * B-spline: scipy.ndimage's spline_filter / map_coordinates (order 3, mode 'mirror', which
    reflects about the edge samples as ITK's mirror boundary does); interpolation of the voxels
    at integer indices.
A plausible slip (taps shifted by one, a radius-3 window, rounding half to even) is shown to
fail the same pin."""
import numpy as np
import pytest
from scipy import ndimage


def _vol(shape=(9, 11, 13), seed=5):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, size=shape).astype(np.float32)


def _points(vol, n=300, seed=7, margin=0.0):
    rng = np.random.default_rng(seed)
    Z, Y, X = vol.shape
    hi = np.array([X - 1, Y - 1, Z - 1], np.float64) - margin
    return (margin + rng.uniform(0.0, 1.0, size=(n, 3)) * (hi - margin)).astype(np.float32)


def test_nn_is_the_voxel_at_the_rounded_index(oracle):
    vol = _vol()
    p = _points(vol)
    frac = p - np.floor(p)
    p = p[np.all(np.abs(frac - 0.5) > 1e-3, axis=1)]
    got = oracle.itk_interp_points(vol, p, 1)
    ix = np.floor(p.astype(np.float64) + 0.5).astype(np.int64)
    want = vol[ix[:, 2], ix[:, 1], ix[:, 0]].astype(np.float64)
    assert np.array_equal(got, want)


def test_nn_rounds_half_integers_up(oracle):
    X = 8
    vol = np.broadcast_to(np.arange(X, dtype=np.float32), (4, 4, X)).copy()  # value = x index
    for x, want in [(0.5, 1), (1.5, 2), (2.5, 3), (3.5, 4), (2.4999998, 2), (2.5000002, 3), (0.0, 0)]:
        got = oracle.itk_interp_points(vol, [[x, 1.0, 1.0]], 1)[0]
        assert got == want, (x, got, want)
    # round-half-to-even would give 2 at 2.5: the pin tells the two apart
    assert np.rint(2.5) == 2.0


def _np_lanczos4(vol, p, radius=4, shift=0):
    """Lanczos-windowed sinc interpolation from np.sinc: weight(t) = sinc(t) sinc(t / radius) at
    t = x - tap over taps floor(x) - radius + 1 + shift .. floor(x) + radius + shift, zero
    outside the volume."""
    pad = radius + 2
    v = np.pad(vol.astype(np.float64), pad)
    out = np.empty(len(p))
    for n, q in enumerate(p.astype(np.float64)):
        b = np.floor(q).astype(np.int64)
        taps = [np.arange(b[k] - radius + 1 + shift, b[k] + radius + 1 + shift) for k in range(3)]
        w = [np.sinc(q[k] - taps[k]) * np.sinc((q[k] - taps[k]) / radius) for k in range(3)]
        blk = v[np.ix_(taps[2] + pad, taps[1] + pad, taps[0] + pad)]
        out[n] = np.einsum("zyx,z,y,x->", blk, w[2], w[1], w[0])
    return out


def test_sinc_matches_numpy_lanczos4(oracle):
    vol = _vol()
    p = _points(vol, n=200)
    got = oracle.itk_interp_points(vol, p, 2)
    want = _np_lanczos4(vol, p)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    # plausible slips fail the same pin
    assert np.abs(_np_lanczos4(vol, p, shift=-1) - got).max() > 1e-3
    assert np.abs(_np_lanczos4(vol, p, radius=3) - got).max() > 1e-3


def test_sinc_at_integer_indices_is_the_voxel(oracle):
    vol = _vol()
    Z, Y, X = vol.shape
    rng = np.random.default_rng(3)
    idx = np.stack([rng.integers(0, X, 50), rng.integers(0, Y, 50), rng.integers(0, Z, 50)], 1)
    got = oracle.itk_interp_points(vol, idx.astype(np.float32), 2)
    assert np.array_equal(got, vol[idx[:, 2], idx[:, 1], idx[:, 0]].astype(np.float64))


def test_sinc_near_the_boundary_sees_zeros(oracle):
    # a constant volume: far inside, the Lanczos-4 weights sum to ~1; half a voxel inside the
    # first face three taps (all with negative weights) lie outside and read 0, so the value
    # overshoots
    vol = np.ones((20, 20, 20), np.float32)
    p = np.array([[9.5, 9.5, 9.5], [0.5, 9.5, 9.5]], np.float32)
    got = oracle.itk_interp_points(vol, p, 2)
    want = _np_lanczos4(vol, p)
    np.testing.assert_allclose(got, want, rtol=1e-12)
    assert abs(got[0] - 1.0) < 0.02 and got[1] > 1.1


def test_bspline_coefficients_match_scipy(oracle):
    vol = _vol(shape=(12, 17, 23))
    got = oracle.bspline_coefficients(vol)
    want = ndimage.spline_filter(vol.astype(np.float64), order=3, mode="mirror")
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("shape", [(9, 11, 13), (24, 30, 40)])
def test_bspline_matches_scipy_map_coordinates(oracle, shape):
    vol = _vol(shape=shape)
    p = _points(vol, n=300)
    got = oracle.itk_interp_points(vol, p, 3)
    want = ndimage.map_coordinates(vol.astype(np.float64), [p[:, 2], p[:, 1], p[:, 0]], order=3, mode="mirror",
                                   prefilter=True)
    np.testing.assert_allclose(got, want, rtol=1e-8, atol=1e-9)
    # taps shifted by one index fail the same pin
    shifted = ndimage.map_coordinates(vol.astype(np.float64), [p[:, 2], p[:, 1], p[:, 0] + 1.0], order=3,
                                      mode="mirror", prefilter=True)
    assert np.abs(shifted - got).max() > 1e-2


def test_bspline_interpolates_the_voxels(oracle):
    vol = _vol()
    Z, Y, X = vol.shape
    rng = np.random.default_rng(4)
    idx = np.stack([rng.integers(0, X, 50), rng.integers(0, Y, 50), rng.integers(0, Z, 50)], 1)
    got = oracle.itk_interp_points(vol, idx.astype(np.float32), 3)
    np.testing.assert_allclose(got, vol[idx[:, 2], idx[:, 1], idx[:, 0]], rtol=1e-9, atol=1e-9)


def test_drr_interp_zero_is_the_linear_drr_and_others_approximate_it(oracle):
    """The DRR loop with interp = 0 is orc_drr_f32 bit for bit; on a smooth volume every other
    interpolator's DRR stays close to it: 3% of the image maximum for nearest neighbour, 1% for
    sinc and B-spline (they differ only in how a sample is read)."""
    from xreg_amd import synth

    w = synth.config("c1", n_poses=2)
    from xreg_amd import geometry as G

    cam = G.downsample_camera(w.cams[0], 0.25)
    z, y, x = np.mgrid[0:w.vol.shape[0], 0:w.vol.shape[1], 0:w.vol.shape[2]].astype(np.float64)
    smooth = (1.0 + 0.5 * np.cos(2 * np.pi * x / 40.0) * np.cos(2 * np.pi * y / 50.0) *
              np.cos(2 * np.pi * z / 30.0)).astype(np.float32)
    poses = np.asarray(w.poses, np.float32).reshape(-1, 16)
    lin, s0 = oracle.drr_f32(smooth, w.idx_to_phys, [cam], poses, step=w.step)
    lin0, s1 = oracle.drr_f32(smooth, w.idx_to_phys, [cam], poses, step=w.step, interp=0)
    assert np.array_equal(lin, lin0) and s0 == s1
    for interp, bound in ((1, 3e-2), (2, 1e-2), (3, 1e-2)):
        got, s = oracle.drr_f32(smooth, w.idx_to_phys, [cam], poses, step=w.step, interp=interp)
        assert s == s0  # the same rays and samples
        rel = np.abs(got.astype(np.float64) - lin) / max(float(lin.max()), 1e-30)
        assert rel.max() < bound, (interp, rel.max())
