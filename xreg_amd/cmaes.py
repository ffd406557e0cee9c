"""CMA-ES in FP64 with an explicit seed, written from N. Hansen, "The CMA Evolution Strategy:
A Tutorial" (arXiv:1604.00772, 2016), Table 1 defaults with positive recombination weights.

Used the way Intensity2D3DRegiCMAES::run drives its optimiser
(lib/regi/interfaces_2d_3d/xregIntensity2D3DRegiCMAES.cpp:75-261): start at a zero delta with a
per-parameter standard deviation, sample lambda candidates per generation, resample any
candidate outside the symmetric box bounds, minimise the objective, stop on the tolerance
tests or an iteration cap. Unlike the vendored cmaes.c (time-seeded when seed = 0) every run
here is reproducible.
"""
from __future__ import annotations

import math

import numpy as np


class CMAES:
    def __init__(self, x0, stddevs, lam: int, seed: int, bounds=None, tol_fun: float = 1e-12,
                 tol_x: float = 1e-11, max_resample: int = 10000):
        x0 = np.asarray(x0, dtype=np.float64)
        sd = np.asarray(stddevs, dtype=np.float64)
        n = x0.size
        self.n = n
        self.lam = int(lam)
        self.rng = np.random.default_rng(seed)
        self.bounds = None if bounds is None else np.asarray(bounds, dtype=np.float64)
        self.mu = self.lam // 2
        w = math.log((self.lam + 1) / 2.0) - np.log(np.arange(1, self.mu + 1))
        self.w = w / w.sum()
        self.mueff = 1.0 / float(np.sum(self.w ** 2))
        self.cs = (self.mueff + 2.0) / (n + self.mueff + 5.0)
        self.ds = 1.0 + 2.0 * max(0.0, math.sqrt((self.mueff - 1.0) / (n + 1.0)) - 1.0) + self.cs
        self.cc = (4.0 + self.mueff / n) / (n + 4.0 + 2.0 * self.mueff / n)
        self.c1 = 2.0 / ((n + 1.3) ** 2 + self.mueff)
        self.cmu = min(1.0 - self.c1, 2.0 * (self.mueff - 2.0 + 1.0 / self.mueff) / ((n + 2.0) ** 2 + self.mueff))
        self.chi_n = math.sqrt(n) * (1.0 - 1.0 / (4.0 * n) + 1.0 / (21.0 * n * n))
        # initial distribution: sigma = RMS of the per-parameter deviations, C diagonal
        self.sigma = float(math.sqrt(np.mean(sd ** 2)))
        self.m = x0.copy()
        self.C = np.diag((sd / self.sigma) ** 2)
        self.B = np.eye(n)
        self.D = sd / self.sigma
        self.pc = np.zeros(n)
        self.ps = np.zeros(n)
        self.gen = 0
        self.tol_fun = tol_fun
        self.tol_x = tol_x
        self.max_resample = max_resample
        self.best_f = math.inf
        self.best_x = x0.copy()
        self.history = []
        self.num_rejects = 0
        self._y = None

    def _sample_one(self):
        z = self.rng.standard_normal(self.n)
        y = self.B @ (self.D * z)
        return y

    def ask(self) -> np.ndarray:
        """lambda candidates (rows); out-of-bounds candidates are re-sampled
        (cmaes_ReSampleSingle, xregIntensity2D3DRegiCMAES.cpp:162-198)."""
        ys = np.empty((self.lam, self.n))
        for k in range(self.lam):
            y = self._sample_one()
            if self.bounds is not None:
                tries = 0
                while np.any(np.abs(self.m + self.sigma * y) > self.bounds):
                    self.num_rejects += 1
                    tries += 1
                    if tries >= self.max_resample:
                        y = np.clip(self.m + self.sigma * y, -self.bounds, self.bounds) - self.m
                        y /= self.sigma
                        break
                    y = self._sample_one()
            ys[k] = y
        self._y = ys
        return self.m[None, :] + self.sigma * ys

    def tell(self, f) -> None:
        f = np.asarray(f, dtype=np.float64)
        assert f.shape == (self.lam,) and self._y is not None
        order = np.argsort(f, kind="stable")
        if f[order[0]] < self.best_f:
            self.best_f = float(f[order[0]])
            self.best_x = self.m + self.sigma * self._y[order[0]]
        self.history.append(float(f[order[0]]))
        y_sel = self._y[order[: self.mu]]
        yw = self.w @ y_sel
        self.m = self.m + self.sigma * yw
        # C^{-1/2} yw = B D^{-1} B^T yw
        cinv_yw = self.B @ ((self.B.T @ yw) / self.D)
        self.ps = (1.0 - self.cs) * self.ps + math.sqrt(self.cs * (2.0 - self.cs) * self.mueff) * cinv_yw
        self.gen += 1
        ps_norm = float(np.linalg.norm(self.ps))
        hs = ps_norm / math.sqrt(1.0 - (1.0 - self.cs) ** (2 * self.gen)) < (1.4 + 2.0 / (self.n + 1.0)) * self.chi_n
        self.pc = (1.0 - self.cc) * self.pc + (math.sqrt(self.cc * (2.0 - self.cc) * self.mueff) * yw if hs else 0.0)
        dh = (1.0 - float(hs)) * self.cc * (2.0 - self.cc)
        rank_mu = (y_sel.T * self.w) @ y_sel
        self.C = ((1.0 + self.c1 * dh - self.c1 - self.cmu) * self.C + self.c1 * np.outer(self.pc, self.pc)
                  + self.cmu * rank_mu)
        self.sigma *= math.exp((self.cs / self.ds) * (ps_norm / self.chi_n - 1.0))
        self.C = 0.5 * (self.C + self.C.T)
        evals, evecs = np.linalg.eigh(self.C)
        self.D = np.sqrt(np.maximum(evals, 1e-300))
        self.B = evecs
        self._y = None

    def stop(self) -> bool:
        if len(self.history) >= 10:
            recent = self.history[-10:]
            if max(recent) - min(recent) < self.tol_fun:
                return True
        if self.sigma * float(np.max(self.D)) < self.tol_x:
            return True
        return False
