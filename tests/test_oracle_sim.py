"""Pins the oracle's similarity metrics against closed forms, scipy's Sobel/correlate with
mirror (REFLECT_101) borders, and brute-force per-patch Pearson correlation."""
import numpy as np
import pytest
from scipy import ndimage


def _img(seed, shape=(30, 34)):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:shape[0], 0:shape[1]]
    return (np.sin(x / 3.0) + np.cos(y / 4.0) + 0.2 * rng.normal(size=shape)).astype(np.float32)


def test_ncc_affine_closed_forms(oracle):
    f = _img(1)
    n = f.size
    pos = oracle.ncc_f32(f, (2.5 * f + 7.0)[None])[0]
    neg = oracle.ncc_f32(f, (-0.5 * f + 1.0)[None])[0]
    # the reference divides by n while its std uses n-1: ncc = (n-1)/n for a > 0
    assert abs(pos - (1 - (n - 1) / n) / 2) < 2e-6
    assert abs(neg - (1 + (n - 1) / n) / 2) < 2e-6
    const = oracle.ncc_f32(f, np.full_like(f, 3.0)[None])[0]
    assert const == 0.5


def test_ncc_matches_pearson(oracle):
    f, m = _img(1), _img(2)
    r = np.corrcoef(f.ravel().astype(np.float64), m.ravel().astype(np.float64))[0, 1]
    n = f.size
    want = (1 - r * (n - 1) / n) / 2
    assert abs(oracle.ncc_f64(f, m[None])[0] - want) < 1e-12
    assert abs(oracle.ncc_f32(f, m[None])[0] - want) < 2e-6
    # masked: the same over the mask only
    mask = np.zeros(f.shape, np.uint8)
    mask[4:20, 6:30] = 1
    sel = mask.astype(bool)
    r = np.corrcoef(f[sel].astype(np.float64), m[sel].astype(np.float64))[0, 1]
    k = int(sel.sum())
    assert abs(oracle.ncc_f32(f, m[None], mask=mask)[0] - (1 - r * (k - 1) / k) / 2) < 2e-6


def test_sobel_matches_scipy_mirror(oracle):
    img = _img(3).astype(np.float64)
    gx, gy = oracle.sobel_f64(img, 0)
    assert np.allclose(gx, ndimage.sobel(img, axis=1, mode="mirror"), atol=1e-12)
    assert np.allclose(gy, ndimage.sobel(img, axis=0, mode="mirror"), atol=1e-12)
    fx, fy = oracle.sobel_f32(img.astype(np.float32), 0)
    assert np.allclose(fx, gx, atol=1e-5) and np.allclose(fy, gy, atol=1e-5)


@pytest.mark.parametrize("k", [3, 5, 7])
def test_blur_fixed_table_then_sobel(oracle, k):
    img = _img(4).astype(np.float64)
    tab = {3: [1, 2, 1], 5: [1, 4, 6, 4, 1], 7: [1, 7.0 / 2, 7, 9, 7, 7.0 / 2, 1]}[k]
    w = np.array(tab, float)
    w = w / w.sum()
    sm = ndimage.correlate1d(ndimage.correlate1d(img, w, axis=1, mode="mirror"), w, axis=0, mode="mirror")
    gx, gy = oracle.sobel_f64(img, k)
    assert np.allclose(gx, ndimage.sobel(sm, axis=1, mode="mirror"), atol=1e-12)
    assert np.allclose(gy, ndimage.sobel(sm, axis=0, mode="mirror"), atol=1e-12)
    fx, fy = oracle.sobel_f32(img.astype(np.float32), k)
    assert np.allclose(fx, gx, atol=2e-5)


def test_sobel_ramp_closed_form(oracle):
    a, b = 0.75, -0.5
    y, x = np.mgrid[0:12, 0:15]
    img = (a * x + b * y).astype(np.float32)
    gx, gy = oracle.sobel_f32(img, 0)
    assert np.allclose(gx[:, 1:-1], 8 * a) and np.allclose(gy[1:-1, :], 8 * b)
    # REFLECT_101 mirrors about the edge pixel: the derivative across the border vanishes
    assert np.allclose(gx[:, 0], 0) and np.allclose(gx[:, -1], 0)
    assert np.allclose(gy[0, :], 0) and np.allclose(gy[-1, :], 0)


def _patch_brute(f, m, r, s=1):
    H, W = f.shape
    d = 2 * r + 1
    N = d * d
    vals = []
    for y0 in range(0, H - d + 1, s):
        for x0 in range(0, W - d + 1, s):
            a = f[y0:y0 + d, x0:x0 + d].ravel().astype(np.float64)
            b = m[y0:y0 + d, x0:x0 + d].ravel().astype(np.float64)
            sa = max(1e-6, a.std(ddof=1))
            sb = max(1e-6, b.std(ddof=1))
            rho = ((a - a.mean()) * (b - b.mean())).sum() / ((N - 1) * sa * sb)
            vals.append(1 - rho * (N - 1) / N)
    return float(np.mean(vals)), len(vals)


@pytest.mark.parametrize("r,s", [(5, 1), (2, 1), (3, 2)])
def test_patch_ncc_brute_force(oracle, r, s):
    f, m = _img(5, (27, 31)), _img(6, (27, 31))
    want, count = _patch_brute(f, m, r, s)
    d = 2 * r + 1
    assert count == ((27 - d) // s + 1) * ((31 - d) // s + 1)
    assert abs(oracle.patch_ncc_f64(f, m[None], r, s)[0] - want) < 1e-12
    assert abs(oracle.patch_ncc_f32(f, m[None], r, s)[0] - want) < 5e-6
    # identical images: every patch scores 1 - (N-1)/N
    N = d * d
    assert abs(oracle.patch_ncc_f32(f, f[None], r, s)[0] - 1.0 / N) < 5e-6


def test_patch_grad_ncc_brute_force(oracle):
    f, m = _img(7, (29, 33)), _img(8, (29, 33))
    fs = ndimage.correlate1d(ndimage.correlate1d(f.astype(np.float64), np.array([1, 4, 6, 4, 1]) / 16.0, axis=1,
                                                 mode="mirror"), np.array([1, 4, 6, 4, 1]) / 16.0, axis=0,
                             mode="mirror")
    ms = ndimage.correlate1d(ndimage.correlate1d(m.astype(np.float64), np.array([1, 4, 6, 4, 1]) / 16.0, axis=1,
                                                 mode="mirror"), np.array([1, 4, 6, 4, 1]) / 16.0, axis=0,
                             mode="mirror")
    x, _ = _patch_brute(ndimage.sobel(fs, 1, mode="mirror"), ndimage.sobel(ms, 1, mode="mirror"), 5)
    y, _ = _patch_brute(ndimage.sobel(fs, 0, mode="mirror"), ndimage.sobel(ms, 0, mode="mirror"), 5)
    want = 0.5 * (x + y)
    assert abs(oracle.patch_grad_ncc_f64(f, m[None], 5, 5)[0] - want) < 1e-12
    assert abs(oracle.patch_grad_ncc_f32(f, m[None], 5, 5)[0] - want) < 5e-6


def test_grad_ncc_brute_force(oracle):
    f, m = _img(9), _img(10)
    gxf, gyf = ndimage.sobel(f.astype(np.float64), 1, mode="mirror"), ndimage.sobel(f.astype(np.float64), 0,
                                                                                      mode="mirror")
    gxm, gym = ndimage.sobel(m.astype(np.float64), 1, mode="mirror"), ndimage.sobel(m.astype(np.float64), 0,
                                                                                      mode="mirror")
    n = f.size

    def ncc(a, b):
        r = np.corrcoef(a.ravel(), b.ravel())[0, 1]
        return (1 - r * (n - 1) / n) / 2

    want = 0.5 * (ncc(gxf, gxm) + ncc(gyf, gym))
    assert abs(oracle.grad_ncc_f64(f, m[None], 0)[0] - want) < 1e-12
    assert abs(oracle.grad_ncc_f32(f, m[None], 0)[0] - want) < 2e-6


def _patch_masked_brute(fixed, mov, mask, r=2, s=1):
    """Masked Patch-NCC with the reference's default options, straight from the definition,
    in float64: weights = unmasked fraction normalised to sum 1, patches of weight <= 1e-6
    skipped, statistics over the whole patch, correlation over unmasked pixels only."""
    d = 2 * r + 1
    rows, cols = fixed.shape
    N = d * d
    centres = [(y, x) for y in range(r, rows - r, s) for x in range(r, cols - r, s)]
    wr = np.array([mask[y - r:y + r + 1, x - r:x + r + 1].astype(bool).sum() / N for y, x in centres])
    w = wr / wr.sum()
    vals = []
    for (y, x), wp in zip(centres, w):
        if not abs(wp) > 1e-6:
            vals.append(0.0)
            continue
        F = fixed[y - r:y + r + 1, x - r:x + r + 1].astype(np.float64)
        M = mov[y - r:y + r + 1, x - r:x + r + 1].astype(np.float64)
        K = mask[y - r:y + r + 1, x - r:x + r + 1].astype(bool)
        sf = max(1e-6, np.sqrt(((F - F.mean()) ** 2).sum() / (N - 1)))
        sm = max(1e-6, np.sqrt(((M - M.mean()) ** 2).sum() / (N - 1)))
        acc = (((M - M.mean()) / sm) * ((F - F.mean()) / (sf * N)))[K].sum()
        vals.append(wp * (1.0 - acc))
    return float(np.sum(vals) / w.sum())


def test_patch_ncc_mask_matches_definition(oracle):
    rng = np.random.default_rng(21)
    fixed = rng.random((14, 17))
    mov = (0.7 * fixed + 0.3 * rng.random((14, 17)))
    mask = np.ones((14, 17), np.uint8)
    mask[:6, :8] = 0          # whole patches masked out (weight 0, skipped)
    mask[9:, 11:13] = 0       # partial patches
    want = _patch_masked_brute(fixed, mov, mask, r=2, s=1)
    got64 = oracle.patch_ncc_f64(fixed, mov[None], radius=2, stride=1, mask=mask)[0]
    got32 = oracle.patch_ncc_f32(fixed, mov[None], radius=2, stride=1, mask=mask)[0]
    assert abs(got64 - want) < 1e-12
    assert abs(got32 - want) < 2e-6
    # a plausible mistake: statistics over unmasked pixels only changes the value
    d, r = 5, 2
    centres = [(y, x) for y in range(r, 14 - r) for x in range(r, 17 - r)]
    assert len(centres) == (14 - 4) * (17 - 4)
    no_mask = oracle.patch_ncc_f64(fixed, mov[None], radius=2, stride=1)[0]
    assert abs(no_mask - want) > 1e-4


def test_patch_ncc_all_ones_mask_is_unmasked(oracle):
    rng = np.random.default_rng(22)
    fixed = rng.random((20, 20)).astype(np.float32)
    mov = rng.random((3, 20, 20)).astype(np.float32)
    ones = np.ones((20, 20), np.uint8)
    a = oracle.patch_ncc_f32(fixed, mov, radius=3, stride=2, mask=ones)
    b = oracle.patch_ncc_f32(fixed, mov, radius=3, stride=2)
    np.testing.assert_allclose(a, b, rtol=0, atol=2e-6)
    a = oracle.patch_grad_ncc_f64(fixed, mov, smooth_ksize=3, radius=3, stride=1, mask=ones)
    b = oracle.patch_grad_ncc_f64(fixed, mov, smooth_ksize=3, radius=3, stride=1)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)
