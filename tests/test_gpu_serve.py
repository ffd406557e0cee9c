"""The continuous-batching registration server: CMA-ES registrations converge, and a job's
results are the same alone and batched with other jobs."""
import numpy as np
import pytest

from tests.helpers import poses_f32

pytestmark = pytest.mark.gpu


def _jobs(oracle, ctx, n_jobs, seed0=0):
    from xreg_amd import _xrh as X
    from xreg_amd import geometry as G
    from xreg_amd import serve, synth

    specs, vols = [], {}
    for name in ("c1", "c2"):
        w = synth.config(name, n_poses=1, small=True)
        vols[name] = (w, X.Volume(ctx, w.vol, w.idx_to_phys))
    rng = np.random.default_rng(seed0)
    for j in range(n_jobs):
        name = ("c1", "c2")[j % 2]
        w, _ = vols[name]
        if j % 4 == 3:  # three views 30 degrees apart
            c0 = w.cams[0]
            cams = synth.multiview_cams(1020.0, c0.num_det_rows, c0.num_det_cols, c0.det_col_spacing, 3, 30.0)
        else:
            cams = [w.cams[0] if j % 3 else G.downsample_camera(w.cams[0], 0.5)]
        fixed = []
        for v, cam in enumerate(cams):
            g, _ = oracle.drr_f32(w.vol, w.idx_to_phys, [cam], poses_f32(w.gt[None]), step=w.step)
            fixed.append(synth.noisy_fixed(g[0], seed=100 + 3 * j + v))
        r = np.deg2rad(rng.uniform(-3, 3, 3))
        t = rng.uniform(-5, 5, 3)
        init = G.perturbed_pose(w.gt, w.center, r, t)
        specs.append(serve.JobSpec(job_id=j, arrival=0.002 * j, volume=name, cams=cams, fixed=fixed,
                                   init_pose=init, center=w.center, step=w.step, lam=16 + 4 * (j % 3),
                                   max_gens=12, seed=1000 + j,
                                   sigma=(np.deg2rad(3.0),) * 3 + (5.0,) * 3))
    return specs, {k: v[1] for k, v in vols.items()}, vols


def test_alone_equals_mixed(ctx, oracle):
    from xreg_amd import serve

    print("case: serve alone-vs-mixed 5 jobs", flush=True)
    specs, vols, _ = _jobs(oracle, ctx, 5)
    srv = serve.RegistrationServer(ctx, vols, hbm_budget=8 << 30, max_items=256, max_views=16, max_det=96)
    alone = {s.job_id: srv.run([s], realtime=False)[0] for s in specs}
    mixed = {r.job_id: r for r in srv.run(specs, realtime=False)}
    srv.close()
    for j, a in alone.items():
        m = mixed[j]
        assert a.history == m.history, j
        assert np.array_equal(a.best_params, m.best_params), j
        assert np.array_equal(a.params, m.params), j
    for v in vols.values():
        v.close()


def test_registration_improves_pose(ctx, oracle):
    from xreg_amd import geometry as G
    from xreg_amd import serve

    print("case: serve convergence", flush=True)
    specs, vols, wl = _jobs(oracle, ctx, 2, seed0=7)
    for s in specs:
        s.max_gens = 40
        s.lam = 24
    srv = serve.RegistrationServer(ctx, vols, hbm_budget=8 << 30, max_items=256, max_views=16, max_det=96)
    res = {r.job_id: r for r in srv.run(specs, realtime=True)}
    srv.close()
    for s in specs:
        r = res[s.job_id]
        w = wl[s.volume][0]
        # mean projection distance of points about the volume centre (the in-plane error a
        # single view constrains)
        ext = np.array(w.vol.shape[::-1]) * w.spacing
        pts = [w.center + 0.25 * ext * np.array(c) for c in [(-1, -1, -1), (1, 1, 1), (1, -1, 1), (-1, 1, -1),
                                                              (0, 0, 0)]]
        err0 = G.mean_projection_distance(s.init_pose, w.gt, pts, 1020.0)
        err1 = G.mean_projection_distance(r.pose, w.gt, pts, 1020.0)  # the CMA-ES mean
        print(f"job {s.job_id}: mean projection distance {err0:.2f} -> {err1:.2f} mm, sim {r.history[0]:.4f} -> "
              f"{r.best_sim:.4f}, latency {r.latency:.3f}s")
        assert r.best_sim < r.history[0]
        assert err1 < 0.5 * err0
    for v in vols.values():
        v.close()


def test_volumes_of_different_storage_in_one_server(ctx, oracle):
    """Jobs on an FP32 volume and on an FP16 volume served together: each submission holds one
    storage type, and every job's result equals its run alone."""
    from xreg_amd import _xrh as X
    from xreg_amd import serve

    print("case: serve FP32 + FP16 volumes", flush=True)
    specs, vols, wl = _jobs(oracle, ctx, 4)
    w2 = wl["c2"][0]
    vols["c2"].close()
    vols["c2"] = X.Volume(ctx, w2.vol, w2.idx_to_phys, storage=X.XRH_F16)
    assert (vols["c1"].storage, vols["c2"].storage) == (X.XRH_F32, X.XRH_F16)
    assert vols["c1"].layout != vols["c2"].layout
    srv = serve.RegistrationServer(ctx, vols, hbm_budget=8 << 30, max_items=256, max_views=16, max_det=96)
    mixed = {r.job_id: r for r in srv.run(specs, realtime=False)}
    assert len(mixed) == 4
    for s in specs:
        alone = srv.run([s], realtime=False)[0]
        assert alone.history == mixed[s.job_id].history, s.job_id
    srv.close()
    for v in vols.values():
        v.close()
