"""CMA-ES driver (xreg_amd/cmaes.py): convergence on textbook functions, bounds, determinism."""
import numpy as np

from xreg_amd.cmaes import CMAES


def _run(f, n=6, lam=12, seed=3, gens=400, bounds=None, sd=0.5, x0=None):
    es = CMAES(np.zeros(n) if x0 is None else x0, np.full(n, sd), lam, seed, bounds=bounds)
    for _ in range(gens):
        X = es.ask()
        es.tell(np.array([f(x) for x in X]))
        if es.stop():
            break
    return es


def test_sphere_converges():
    target = np.array([0.3, -0.2, 0.1, 0.05, -0.4, 0.25])
    es = _run(lambda x: float(np.sum((x - target) ** 2)))
    assert np.linalg.norm(es.m - target) < 1e-5


def test_rotated_ellipsoid_converges():
    rng = np.random.default_rng(0)
    Q, _ = np.linalg.qr(rng.normal(size=(6, 6)))
    scales = 10.0 ** np.linspace(0, 3, 6)
    es = _run(lambda x: float(np.sum(scales * (Q @ x - 0.1) ** 2)), gens=1500, lam=16)
    assert es.best_f < 1e-8


def test_bounds_are_respected_and_resampled():
    b = np.full(6, 0.2)
    es = CMAES(np.zeros(6), np.full(6, 0.5), 20, seed=1, bounds=b)
    X = es.ask()
    assert np.all(np.abs(X) <= b + 1e-12)
    assert es.num_rejects > 0


def test_same_seed_same_trajectory():
    f = lambda x: float(np.sum(np.cos(3 * x) + x ** 2))  # noqa: E731
    a = _run(f, gens=30, seed=42)
    b = _run(f, gens=30, seed=42)
    c = _run(f, gens=30, seed=43)
    assert np.array_equal(a.m, b.m) and a.history == b.history
    assert not np.array_equal(a.m, c.m)


def test_population_and_weights():
    es = CMAES(np.zeros(6), np.ones(6), 100, seed=0)
    assert es.mu == 50 and abs(es.w.sum() - 1) < 1e-12 and np.all(np.diff(es.w) < 0)
    assert es.ask().shape == (100, 6)
