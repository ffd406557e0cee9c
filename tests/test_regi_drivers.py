"""Registration drivers on the CPU with a closed-form objective (no GPU): the exhaustive
search's batching and first-minimum rule, the refiner's BOBYQA-shaped call pattern and
convergence, CMA-ES convergence, and CombineMean's float arithmetic."""
import numpy as np
import pytest

from xreg_amd import geometry as G
from xreg_amd import regi


class QuadObj:
    """f(x) = sum(a * (x - x*)^2) as float32; records every call's size."""

    def __init__(self, xstar, a, max_batch=16):
        self.xstar, self.a = np.asarray(xstar, float), np.asarray(a, float)
        self.max_batch = max_batch
        self.evals, self.calls, self.sizes = 0, 0, []
        self.param = regi.PoseParam(np.eye(4), np.zeros(3))

    def __call__(self, X):
        X = np.atleast_2d(X)
        out = []
        for s in range(0, len(X), self.max_batch):
            c = X[s:s + self.max_batch]
            self.sizes.append(len(c))
            self.calls += 1
            out.append(np.sum(self.a * (c - self.xstar) ** 2, axis=1).astype(np.float32))
        self.evals += len(X)
        return np.concatenate(out)


def test_exhaustive_first_minimum_and_batches():
    obj = QuadObj(np.zeros(6), np.ones(6), max_batch=16)
    grid = regi.mesh_grid([[-1.0, 0.0, 1.0], [0.0], [0.0], [0.0], [0.0], [-1.0, 1.0]])
    # (-1,..,-1) (−1,..,1) (0,..,-1) (0,..,1) (1,..,-1) (1,..,1): four equal minima at x0 = 0
    r = regi.run_exhaustive(obj, grid, batch=4)
    assert obj.sizes == [4, 2]
    assert r.extra["best_index"] == 2  # the first of the equal minima
    assert np.array_equal(r.x, grid[2])
    assert len(r.extra["all_sim_vals"]) == 6


def test_mesh_grid_order_first_parameter_most_significant():
    g = regi.mesh_grid([[0, 1], [10, 20, 30], [5], [6], [7], [8]])
    assert g.shape == (6, 6)
    assert list(g[:, 0]) == [0, 0, 0, 1, 1, 1] and list(g[:, 1]) == [10, 20, 30, 10, 20, 30]


def test_refiner_call_pattern_and_convergence():
    xstar = np.array([0.02, -0.03, 0.01, 1.5, -2.0, 0.7])
    obj = QuadObj(xstar, [400.0, 300.0, 500.0, 1.0, 2.0, 1.5])
    r = regi.run_local_refine(obj, np.zeros(6), rho_beg=[0.05] * 3 + [2.0] * 3, rho_end=1e-4, max_evals=150)
    assert obj.sizes[:13] == [1] * 13 and all(s == 1 for s in obj.sizes)  # single-pose evaluations
    assert r.evals <= 150
    assert np.allclose(r.x, xstar, atol=1e-2 * np.array([0.05] * 3 + [2.0] * 3))
    assert r.fun < 1e-4


def test_refiner_respects_bounds():
    obj = QuadObj(np.full(6, 5.0), np.ones(6))
    bounds = np.array([[-1.0, 1.0]] * 6)
    r = regi.run_local_refine(obj, np.zeros(6), rho_beg=[0.5] * 6, max_evals=100, bounds=bounds)
    assert np.all(r.x <= 1.0 + 1e-12) and np.all(r.x >= -1.0 - 1e-12)
    assert np.allclose(r.x, 1.0, atol=1e-3)


def test_cmaes_driver_converges():
    xstar = np.array([0.05, -0.02, 0.03, 3.0, -4.0, 2.0])
    obj = QuadObj(xstar, [100.0] * 3 + [1.0] * 3, max_batch=64)
    r = regi.run_cmaes(obj, lam=20, sigma=[0.1] * 3 + [5.0] * 3, max_gens=200, seed=1)
    assert all(s == 20 for s in obj.sizes)  # one submission per generation
    assert np.allclose(r.x, xstar, atol=1e-3)


def test_cmaes_estimate_is_the_mean_and_debug_record(tmp_path):
    """The estimate is the distribution mean after the last generation (as the reference); the
    debug record holds one mean and one best-so-far value per generation, non-increasing."""
    xstar = np.array([0.02, -0.01, 0.03, 1.0, -2.0, 0.5])
    obj = QuadObj(xstar, [100.0] * 3 + [1.0] * 3, max_batch=64)
    r = regi.run_cmaes(obj, lam=16, sigma=[0.05] * 3 + [2.0] * 3, max_gens=30, seed=4)
    means, best = r.extra["iter_means"], r.extra["iter_best"]
    assert means.shape == (len(best), 6) and np.array_equal(r.x, means[-1])
    assert np.all(np.diff(best) <= 0) and best[-1] == r.fun
    np.testing.assert_allclose(r.pose, obj.param(r.x), atol=0)
    h5io = pytest.importorskip("xreg_amd.h5io")
    if not h5io.available():
        pytest.skip("HDF5 C library not available")
    rec = regi.debug_results(r, obj.param, vol_index=0, regi_time_secs=1.25)
    h5io.write_regi_debug(tmp_path / "dbg.h5", rec)
    got = h5io.read_regi_debug(tmp_path / "dbg.h5")
    assert np.array_equal(got.iter_vars[0], means.astype(np.float32))
    assert np.array_equal(got.sims, best.astype(np.float32))
    assert np.array_equal(got.final_poses[0], r.pose.astype(np.float32))
    assert got.inter_frames_wrt_vol == [True] and type(got.se3_params).__name__ == "SE3OptVarsLieAlg"


def test_combine_mean_is_float_sum_over_views_divided():
    v = np.array([[0.1, 0.3], [0.2, 0.7], [0.35, 0.05]], np.float32)
    got = regi.combine_mean(v)
    want = np.array([(np.float32(np.float32(np.float32(0.1) + np.float32(0.2)) + np.float32(0.35))) / np.float32(3),
                     (np.float32(np.float32(np.float32(0.3) + np.float32(0.7)) + np.float32(0.05))) / np.float32(3)],
                    np.float32)
    assert got.dtype == np.float32 and np.array_equal(got, want)


def test_pose_param_batch_matches_single():
    init = G.gt_pose(np.array([10.0, 20.0, 30.0]), 510.0)
    p = regi.PoseParam(init, np.array([10.0, 20.0, 30.0]))
    X = np.random.default_rng(0).normal(0, [0.1] * 3 + [5.0] * 3, (8, 6))
    B = p(X)
    for x, b in zip(X, B):
        np.testing.assert_allclose(b, p(x), atol=1e-9)
    np.testing.assert_allclose(p(np.zeros(6)), init, atol=1e-12)


def test_downsample_camera_follows_reference_arithmetic():
    c = G.camera(1020.0, 375, 384, 0.776, 0.776)
    d = G.downsample_camera(c, 0.5)
    assert (d.num_det_rows, d.num_det_cols) == (188, 192)  # lround(187.5) = 188
    K = G.naive_intrins(1020.0, 375, 384, 0.776, 0.776)
    K2 = K.copy()
    for r, k in ((0, 0), (1, 1), (0, 2), (1, 2)):
        K2[r, k] = np.float32(K[r, k] * np.float32(0.5))
    ref = G.camera_ke(K2, np.eye(4, dtype=np.float32), 188, 192, float(np.float32(0.776) / np.float32(0.5)),
                      float(np.float32(0.776) / np.float32(0.5)))
    assert list(d.intrins_inv) == list(ref.intrins_inv) and d.focal_len == ref.focal_len
    e = G.downsample_camera(c, 0.5, force_even_dims=True)
    assert (e.num_det_rows, e.num_det_cols) == (188, 192)
    assert G.lround(2.5) == 3 and G.lround(-2.5) == -3 and G.lround(2.4999) == 2


def test_multi_param_stacks_object_parameters():
    """MultiParam: object o reads its own slice of the optimiser vector; zeros give each
    object's starting pose; one vector gives a list of poses, k vectors a list of (k, 4, 4)."""
    from xreg_amd import geometry as G
    from xreg_amd import regi

    c = np.array([10.0, 20.0, 30.0])
    a0 = G.perturbed_pose(np.eye(4), c, np.deg2rad([5.0, 0, 0]), [1.0, 2.0, 3.0])
    b0 = G.perturbed_pose(np.eye(4), c, np.deg2rad([0, -7.0, 0]), [-4.0, 0.0, 2.0])
    mp = regi.MultiParam([regi.PoseParam(a0, c), regi.PoseParam(b0, c)])
    assert mp.num_params == 12
    pa, pb = mp(np.zeros(12))
    np.testing.assert_allclose(pa, a0, atol=1e-12)
    np.testing.assert_allclose(pb, b0, atol=1e-12)
    x = np.zeros((3, 12))
    x[1, 3] = 5.0     # object A, x translation
    x[2, 9] = -2.0    # object B, x translation
    A, B = mp(x)
    assert A.shape == (3, 4, 4) and B.shape == (3, 4, 4)
    np.testing.assert_allclose(A[2], a0, atol=1e-12)   # B's parameters leave A alone
    np.testing.assert_allclose(B[1], b0, atol=1e-12)
    np.testing.assert_allclose(A[1], regi.PoseParam(a0, c)(x[1, :6]), atol=1e-12)
    np.testing.assert_allclose(B[2], regi.PoseParam(b0, c)(x[2, 6:]), atol=1e-12)
