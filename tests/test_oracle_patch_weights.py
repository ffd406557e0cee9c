"""Pins the oracle's general patch combine (patch weights, weight images, the patches used,
weight-in-combine and mean flags of ImgSimMetric2DPatchCommon / PatchNCCCPU::compute) against
per-patch scores computed from the definition (Pearson correlation per window) in float64."""
import numpy as np
import pytest
from numpy.lib.stride_tricks import sliding_window_view
from scipy import ndimage


def _img(seed, shape):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:shape[0], 0:shape[1]]
    return (np.sin(x / 3.0) + np.cos(y / 4.0) + 0.3 * rng.normal(size=shape)).astype(np.float32)


def _patch_scores(fixed, mov, r, s=1, mask=None):
    """s_p = 1 - sum_U ((m - mu_m) / sd_m) (f - mu_f) / (sd_f N) for every patch, rows of
    centres outer; statistics over the whole patch, the sum over unmasked pixels U."""
    d = 2 * r + 1
    N = d * d
    rows, cols = fixed.shape
    out = []
    for y in range(r, rows - r, s):
        for x in range(r, cols - r, s):
            F = fixed[y - r:y + r + 1, x - r:x + r + 1].astype(np.float64)
            M = mov[y - r:y + r + 1, x - r:x + r + 1].astype(np.float64)
            K = np.ones_like(F, bool) if mask is None else mask[y - r:y + r + 1, x - r:x + r + 1].astype(bool)
            sf = max(1e-6, F.std(ddof=1))
            sm = max(1e-6, M.std(ddof=1))
            out.append(1.0 - ((((M - M.mean()) / sm) * ((F - F.mean()) / (sf * N)))[K]).sum())
    return np.array(out)


@pytest.fixture(scope="module")
def pair():
    f, m = _img(31, (21, 24)), _img(32, (21, 24))
    m = (0.6 * f + 0.4 * m).astype(np.float32)
    return f, m


def test_single_patch_is_its_pearson_score(oracle, pair):
    f, m = pair
    r, d = 2, 5
    s = _patch_scores(f, m, r)
    # the definition is 1 - (N-1)/N Pearson(f, m) over the window
    q = 37
    ppr = 24 - d + 1
    y, x = q // ppr, q % ppr
    rho = np.corrcoef(f[y:y + d, x:x + d].ravel().astype(np.float64), m[y:y + d, x:x + d].ravel())[0, 1]
    assert abs(s[q] - (1 - rho * 24 / 25)) < 1e-12
    for wic in (False, True):
        got = oracle.patch_sim_weighted_f32(2, f, m[None], inds=[q], weight_in_combine=wic, radius=r)[0]
        assert abs(got - s[q]) < 2e-6
        got = oracle.patch_sim_weighted_f64(2, f, m[None], inds=[q], weight_in_combine=wic, radius=r)[0]
        assert abs(got - s[q]) < 1e-12


def test_repeats_sum_and_mean(oracle, pair):
    f, m = pair
    s = _patch_scores(f, m, 2)
    inds = [5, 5, 80, 3]
    want_sum = 2 * s[5] + s[80] + s[3]
    got = oracle.patch_sim_weighted_f64(2, f, m[None], inds=inds, weight_in_combine=False, radius=2)[0]
    assert abs(got - want_sum) < 1e-12
    got = oracle.patch_sim_weighted_f32(2, f, m[None], inds=inds, weight_in_combine=False, radius=2)[0]
    assert abs(got - want_sum) < 1e-5
    got = oracle.patch_sim_weighted_f64(2, f, m[None], inds=inds, weight_in_combine=False, compute_mean=True,
                                        radius=2)[0]
    assert abs(got - want_sum / 4) < 1e-12


def test_weighted_combine_and_divisor(oracle, pair):
    f, m = pair
    for stride in (1, 2):
        s = _patch_scores(f, m, 2, stride)
        rng = np.random.default_rng(stride)
        w = rng.uniform(0.1, 2.0, s.size).astype(np.float32)
        w64 = w.astype(np.float64)
        want = float((w64 * s).sum() / w64.sum())
        got = oracle.patch_sim_weighted_f64(2, f, m[None], weights=w, radius=2, stride=stride)[0]
        assert abs(got - want) < 1e-12
        got = oracle.patch_sim_weighted_f32(2, f, m[None], weights=w, radius=2, stride=stride)[0]
        assert abs(got - want) < 2e-6
        # weights on a subset with repeats: divided by the sum of the used weights
        inds = np.array([4, 9, 9, 1, s.size - 1])
        want = float((w64[inds] * s[inds]).sum() / w64[inds].sum())
        got = oracle.patch_sim_weighted_f64(2, f, m[None], weights=w, inds=inds, radius=2, stride=stride)[0]
        assert abs(got - want) < 1e-12
        # mean of weighted scores: divided by the number used
        want = float((w64[inds] * s[inds]).sum() / inds.size)
        got = oracle.patch_sim_weighted_f64(2, f, m[None], weights=w, inds=inds, compute_mean=True, radius=2,
                                            stride=stride)[0]
        assert abs(got - want) < 1e-12
        # neither: the plain sum, weights ignored
        got = oracle.patch_sim_weighted_f64(2, f, m[None], weights=w, inds=inds, weight_in_combine=False,
                                            radius=2, stride=stride)[0]
        assert abs(got - float(s[inds].sum())) < 1e-12


def test_low_weight_patch_is_skipped_only_when_weighting(oracle, pair):
    f, m = pair
    s = _patch_scores(f, m, 2)
    w = np.ones(s.size, np.float32)
    w[7] = 5e-7
    # alone: its score is skipped (0) and the sum divided by its weight
    assert oracle.patch_sim_weighted_f64(2, f, m[None], weights=w, inds=[7], radius=2)[0] == 0.0
    assert oracle.patch_sim_weighted_f32(2, f, m[None], weights=w, inds=[7], radius=2)[0] == 0.0
    # not weighted in the combine: the weight is ignored
    got = oracle.patch_sim_weighted_f64(2, f, m[None], weights=w, inds=[7], weight_in_combine=False, radius=2)[0]
    assert abs(got - s[7]) < 1e-12
    w[7] = 2e-6
    got = oracle.patch_sim_weighted_f64(2, f, m[None], weights=w, inds=[7], radius=2)[0]
    assert abs(got - s[7]) < 1e-9


def test_mask_restricts_correlation_with_any_weights(oracle, pair):
    f, m = pair
    mask = np.ones(f.shape, np.uint8)
    mask[3:9, 5:14] = 0
    s = _patch_scores(f, m, 2, 1, mask)
    w = np.linspace(0.2, 1.5, s.size).astype(np.float32)
    w64 = w.astype(np.float64)
    want = float((w64 * s).sum() / w64.sum())
    got = oracle.patch_sim_weighted_f64(2, f, m[None], weights=w, mask=mask, radius=2)[0]
    assert abs(got - want) < 1e-12
    # the mask's own weights through the general entry equal the masked default
    wm = oracle.patch_weights_f32(f.shape, radius=2, mask=mask)
    a = oracle.patch_sim_weighted_f32(2, f, m[None], weights=wm, mask=mask, radius=2)
    b = oracle.patch_ncc_f32(f, m[None], radius=2, mask=mask)
    assert np.array_equal(a, b)


def test_grad_metric_uses_the_same_patches_in_x_and_y(oracle, pair):
    f, m = pair
    gxf, gyf = oracle.sobel_f64(f.astype(np.float64), 0)
    gxm, gym = oracle.sobel_f64(m.astype(np.float64), 0)
    # Sobel with the mirror border from scipy (pinned in the similarity tests)
    assert np.allclose(gxf, ndimage.sobel(f.astype(np.float64), 1, mode="mirror"))
    sx = _patch_scores(gxf, gxm, 2)
    sy = _patch_scores(gyf, gym, 2)
    w = np.random.default_rng(3).uniform(0.5, 1.0, sx.size).astype(np.float32)
    inds = [2, 50, 50, 61]
    wu = w[inds].astype(np.float64)
    want = 0.5 * ((wu * sx[inds]).sum() + (wu * sy[inds]).sum()) / wu.sum()
    got = oracle.patch_sim_weighted_f64(3, f, m[None], weights=w, inds=inds, smooth_ksize=0, radius=2)[0]
    assert abs(got - want) < 1e-12
    got = oracle.patch_sim_weighted_f32(3, f, m[None], weights=w, inds=inds, smooth_ksize=0, radius=2)[0]
    assert abs(got - want) < 5e-6
    # unit weights over every patch: the default Patch-Grad-NCC
    a = oracle.patch_sim_weighted_f32(3, f, m[None], smooth_ksize=5, radius=2)
    b = oracle.patch_grad_ncc_f32(f, m[None], smooth_ksize=5, radius=2)
    assert np.array_equal(a, b)


def test_patch_weights_from_weight_image_and_mask(oracle):
    shape, r, s = (17, 20), 2, 2
    ppr = (20 - 5) // s + 1
    ppc = (17 - 5) // s + 1
    assert np.array_equal(oracle.patch_weights_f32(shape, r, s), np.ones(ppr * ppc, np.float32))
    # a weight image that is 1 at one patch centre only: that patch gets all the weight
    wi = np.zeros(shape, np.float32)
    wi[r + 3 * s, r + 4 * s] = 1.0
    w = oracle.patch_weights_f32(shape, r, s, wgt_img=wi)
    want = np.zeros(ppr * ppc, np.float32)
    want[3 * ppr + 4] = 1.0
    assert np.array_equal(w, want)
    # values at the centres, unnormalised; a mask off at a centre zeroes it when used
    wi = np.arange(np.prod(shape), dtype=np.float32).reshape(shape)
    centres = wi[r:17 - r:s, r:20 - r:s].ravel()
    assert np.array_equal(oracle.patch_weights_f32(shape, r, s, wgt_img=wi, normalize=False), centres)
    mask = np.ones(shape, np.uint8)
    mask[r + s, :] = 0
    w = oracle.patch_weights_f32(shape, r, s, wgt_img=wi, mask=mask, normalize=False)
    assert np.all(w[ppr:2 * ppr] == 0) and np.array_equal(w[:ppr], centres[:ppr])
    w = oracle.patch_weights_f32(shape, r, s, wgt_img=wi, mask=mask, use_mask_for_weighting=False,
                                 normalize=False)
    assert np.array_equal(w, centres)
    # normalised: sums to one
    w = oracle.patch_weights_f32(shape, r, s, wgt_img=wi)
    assert abs(float(w.astype(np.float64).sum()) - 1.0) < 1e-6
    # mask only: unmasked fraction of each window, normalised
    mask = (np.random.default_rng(4).random(shape) > 0.3).astype(np.uint8)
    frac = sliding_window_view(mask.astype(np.float64), (5, 5))[::s, ::s].mean(axis=(-1, -2)).ravel()
    w = oracle.patch_weights_f32(shape, r, s, mask=mask)
    np.testing.assert_allclose(w, frac / frac.sum(), rtol=2e-6, atol=0)
    # a mask not used for weighting and no image: unit weights
    assert np.array_equal(oracle.patch_weights_f32(shape, r, s, mask=mask, use_mask_for_weighting=False),
                          np.ones(ppr * ppc, np.float32))


def test_patch_variance_image_matches_numpy(oracle):
    """use_fixed_img_patch_variances_as_wgts: the variance of each patch (over its unmasked
    pixels with a mask; n - 1 denominator, std clamped at 1e-6) at the patch centre."""
    shape, r, s = (19, 23), 2, 2
    img = _img(41, shape)
    img[:6, :9] = 3.0                    # flat patches: std clamped at 1e-6
    mask = (np.random.default_rng(5).random(shape) > 0.25).astype(np.uint8)
    for mk in (None, mask):
        got = oracle.patch_variance_image_f32(img, r, s, mask=mk)
        want = np.zeros(shape, np.float64)
        win = sliding_window_view(img.astype(np.float64), (5, 5))[::s, ::s]
        kw = None if mk is None else sliding_window_view(mk.astype(bool), (5, 5))[::s, ::s]
        for i in range(win.shape[0]):
            for j in range(win.shape[1]):
                v = win[i, j] if kw is None else win[i, j][kw[i, j]]
                var = v.var(ddof=1) if v.size > 1 else 0.0
                want[r + i * s, r + j * s] = max(1e-6, np.sqrt(var)) ** 2
        off = np.ones(shape, bool)
        off[r:shape[0] - r:s, r:shape[1] - r:s] = False
        assert np.all(got[off] == 0)
        np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-13)
        assert got[r, r] == np.float32(1e-6) ** 2 or mk is not None


def _patch_scores_mask_stats(fixed, mov, mask, r, s=1):
    """use_mask_for_patch_stats: means and deviations over the unmasked pixels U of the patch
    (the sum left undivided when |U| <= 1, deviation 0 then, clamped at 1e-6), the fixed patch
    scaled by |U|, the correlation over U; an empty U scores 1."""
    out = []
    rows, cols = fixed.shape
    for y in range(r, rows - r, s):
        for x in range(r, cols - r, s):
            K = mask[y - r:y + r + 1, x - r:x + r + 1].astype(bool)
            F = fixed[y - r:y + r + 1, x - r:x + r + 1].astype(np.float64)[K]
            M = mov[y - r:y + r + 1, x - r:x + r + 1].astype(np.float64)[K]
            n = F.size
            if n == 0:
                out.append(1.0)
                continue
            mf, mm = (F.mean(), M.mean()) if n > 1 else (F.sum(), M.sum())
            sf = max(1e-6, F.std(ddof=1)) if n > 1 else 1e-6
            sm = max(1e-6, M.std(ddof=1)) if n > 1 else 1e-6
            out.append(1.0 - (((M - mm) / sm) * ((F - mf) / (sf * n))).sum())
    return np.array(out)


def test_mask_stats_match_definition(oracle, pair):
    f, m = pair
    mask = (np.random.default_rng(8).random(f.shape) > 0.3).astype(np.uint8)
    mask[:7, :7] = 0          # a patch with no unmasked pixel (score 1)
    mask[9:14, 0:5] = 0
    mask[11, 2] = 1           # a patch with a single unmasked pixel
    for stride in (1, 2):
        s = _patch_scores_mask_stats(f, m, mask, 2, stride)
        assert s[0] == 1.0
        w = oracle.patch_weights_f32(f.shape, 2, stride, mask=mask).astype(np.float64)
        keep = np.abs(w) > 1e-6
        want = float((w * s * keep).sum() / w.sum())
        got = oracle.patch_sim_weighted_f64(2, f, m[None], mask=mask, radius=2, stride=stride, mask_stats=True,
                                            weights=w.astype(np.float32))[0]
        assert abs(got - want) < 1e-9
        # unit weights, not weighted: the plain mean of the scores (the empty patch included)
        got = oracle.patch_sim_weighted_f64(2, f, m[None], mask=mask, radius=2, stride=stride, mask_stats=True,
                                            weight_in_combine=False, compute_mean=True)[0]
        assert abs(got - float(s.mean())) < 1e-12
        got32 = oracle.patch_sim_weighted_f32(2, f, m[None], mask=mask, radius=2, stride=stride, mask_stats=True,
                                              weight_in_combine=False, compute_mean=True)[0]
        assert abs(got32 - float(s.mean())) < 1e-5
    # whole-patch statistics differ (the option matters), and without a mask it changes nothing
    a = oracle.patch_sim_weighted_f64(2, f, m[None], mask=mask, radius=2, mask_stats=True)[0]
    b = oracle.patch_sim_weighted_f64(2, f, m[None], mask=mask, radius=2)[0]
    assert abs(a - b) > 1e-4
    assert np.array_equal(oracle.patch_sim_weighted_f32(2, f, m[None], radius=2, mask_stats=True),
                          oracle.patch_sim_weighted_f32(2, f, m[None], radius=2))


def test_patch_scores_are_the_per_window_scores(oracle, pair):
    """save_all_per_patch_scores: every patch's unweighted score; Patch-Grad-NCC the float mean
    of its x and y scores; masked statistics included."""
    f, m = pair
    got = oracle.patch_scores_f32(2, f, m[None], radius=2, stride=2)[0]
    np.testing.assert_allclose(got, _patch_scores(f, m, 2, 2), rtol=0, atol=2e-6)
    gxf, gyf = oracle.sobel_f64(f.astype(np.float64), 0)
    gxm, gym = oracle.sobel_f64(m.astype(np.float64), 0)
    want = 0.5 * (_patch_scores(gxf, gxm, 2) + _patch_scores(gyf, gym, 2))
    got = oracle.patch_scores_f32(3, f, m[None], smooth_ksize=0, radius=2)[0]
    np.testing.assert_allclose(got, want, rtol=0, atol=5e-6)
    mask = (np.random.default_rng(9).random(f.shape) > 0.3).astype(np.uint8)
    got = oracle.patch_scores_f32(2, f, m[None], radius=2, mask=mask, mask_stats=True)[0]
    np.testing.assert_allclose(got, _patch_scores_mask_stats(f, m, mask, 2), rtol=0, atol=2e-6)
    # their plain mean is the unweighted metric
    sc = oracle.patch_scores_f32(2, f, m[None], radius=2)[0]
    mean = oracle.patch_sim_weighted_f64(2, f, m[None], radius=2, weight_in_combine=False, compute_mean=True)[0]
    assert abs(float(sc.astype(np.float64).mean()) - mean) < 1e-6
