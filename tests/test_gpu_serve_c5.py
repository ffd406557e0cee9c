"""The served registration stream at its own shapes (reduced job count and generations): three
resident phantoms (192^3 @1.5 mm, 256x256x200 @1.0 mm, 512x512x400 @0.5 mm, FP32), detectors
192^2, 384^2 and 768^2, CMA-ES lambda from 20 to 200, one or three views, patch gradient-NCC,
admitted against a 200 GB HBM budget. Every job gives the same result alone and in the mix, and
the first generation's similarity values of one job per phantom match the CPU restatement of
the reference (oracle) within 1e-5; their distance from the FP64 oracle is printed."""
import numpy as np
import pytest

from tests.helpers import SIM_ABS_TOL, poses_f32

pytestmark = pytest.mark.gpu

PHANTOMS = (("ph192", (192, 192, 192), 1.5), ("ph256", (256, 256, 200), 1.0), ("ph512", (512, 512, 400), 0.5))
DET_MM = 192 * 1.552
# (phantom, detector, views, lambda); the first job of each phantom is checked against the oracle
JOBS = (("ph192", 192, 1, 20), ("ph256", 384, 3, 50), ("ph512", 192, 3, 120),
        ("ph512", 768, 1, 200), ("ph256", 768, 1, 80), ("ph192", 384, 3, 160))
GENS = 3
N_ORACLE_POSES = 2


def test_c5_stream_alone_equals_mixed_and_matches_oracle(ctx, oracle):
    from xreg_amd import _xrh as X
    from xreg_amd import geometry as G
    from xreg_amd import serve, synth

    rng = np.random.default_rng(55)
    vols, meta = {}, {}
    for name, size, sp in PHANTOMS:
        print(f"case: c5 phantom {name} {size} @{sp} mm", flush=True)
        vox = synth.phantom(size)
        vols[name] = X.Volume(ctx, vox, G.idx_to_phys(sp))
        meta[name] = (vox, sp, G.vol_center_phys(size, sp))
    buf = ctx.alloc(4 * 768 * 768)
    specs = []
    for j, (name, det, nv, lam) in enumerate(JOBS):
        _, sp, center = meta[name]
        cams = synth.multiview_cams(1020.0, det, det, DET_MM / det, 3, 30.0)[:nv]
        gt = G.perturbed_pose(G.gt_pose(center, 510.0), center, np.deg2rad(rng.uniform(-3, 3, 3)),
                              rng.uniform(-5, 5, 3))
        fixed = []
        for v, cam in enumerate(cams):
            X.line_int(ctx, vols[name], [cam], poses_f32(gt[None]), np.zeros(1, np.int64), sp / 2, buf)
            fixed.append(synth.noisy_fixed(buf.download(np.float32, count=det * det).reshape(det, det),
                                           seed=700 + 3 * j + v))
        init = G.perturbed_pose(gt, center, np.deg2rad(rng.uniform(-5, 5, 3)), rng.uniform(-10, 10, 3))
        specs.append(serve.JobSpec(job_id=j, arrival=0.0, volume=name, cams=cams, fixed=fixed, init_pose=init,
                                   center=center, step=sp / 2, lam=lam, max_gens=GENS, seed=600 + j,
                                   sigma=(np.deg2rad(5.0),) * 3 + (10.0,) * 3))
    buf.free()

    srv = serve.RegistrationServer(ctx, vols, hbm_budget=200 << 30, max_items=2048, max_views=64, max_det=768)
    print("case: c5 jobs alone", flush=True)
    alone = {s.job_id: srv.run([s], realtime=False)[0] for s in specs}
    print("case: c5 jobs mixed", flush=True)
    mixed = {r.job_id: r for r in srv.run(specs, realtime=False)}
    steps_mixed = srv.steps
    srv.close()
    assert steps_mixed > 0
    for j, a in alone.items():
        m = mixed[j]
        assert a.status != "rejected" and m.status != "rejected", j
        assert a.generations == m.generations == GENS, j
        assert a.history == m.history, j
        assert np.array_equal(a.first_sims, m.first_sims), j
        assert np.array_equal(a.params, m.params), j

    checked = set()
    for j, (name, det, nv, lam) in enumerate(JOBS):
        if name in checked:
            continue
        checked.add(name)
        vox, sp, _ = meta[name]
        r = mixed[j]
        pf = poses_f32(r.first_poses[:N_ORACLE_POSES])
        print(f"case: c5 oracle check job {j} ({name}, {det}^2, {nv} views)", flush=True)
        for v, cam in enumerate(specs[j].cams):
            d32, _ = oracle.drr_f32(vox, G.idx_to_phys(sp), [cam], pf, step=sp / 2)
            d64, _ = oracle.drr_f64(vox, G.idx_to_phys(sp), [cam], pf.astype(np.float64), step=sp / 2)
            w32 = oracle.patch_grad_ncc_f32(specs[j].fixed[v], d32, smooth_ksize=5)
            w64 = oracle.patch_grad_ncc_f64(specs[j].fixed[v], d64, smooth_ksize=5)
            got = r.first_sims[v, :N_ORACLE_POSES]
            e32 = np.abs(got - w32).max()
            e64 = np.abs(got.astype(np.float64) - w64).max()
            print(f"  view {v}: |served - fp32 restatement| {e32:.2e}, |served - fp64 oracle| {e64:.2e}", flush=True)
            assert e32 <= SIM_ABS_TOL
            assert np.all(np.abs(got.astype(np.float64) - w64) <= SIM_ABS_TOL + np.abs(w32.astype(np.float64) - w64))
    assert checked == {"ph192", "ph256", "ph512"}
    for v in vols.values():
        v.close()
