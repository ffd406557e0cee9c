"""GPU parity of the line-integral DRR (xrh_line_int) against the CPU restatement of
RayCasterLineIntCPU (oracle/xrh_oracle.c) and the FP64 oracle, through the C-ABI."""
import numpy as np
import pytest

from tests.helpers import DRR_REL_TOL, drr_rel_err, poses_f32

pytestmark = pytest.mark.gpu


def _gpu_drr(ctx, w, poses, layout=None, storage=0, geom=0, kernel=0, store=0, bg=None, default_bg=0.0,
             init=None, cam_for_proj=None, cams=None):
    from xreg_amd import _xrh as X

    cams = cams or w.cams
    rows, cols = cams[0].num_det_rows, cams[0].num_det_cols
    n = poses.shape[0]
    vol = X.Volume(ctx, w.vol, w.idx_to_phys, storage=storage, layout=layout)
    buf = ctx.alloc(4 * n * rows * cols)
    if init is not None:
        buf.upload(init.astype(np.float32))
    bgb = None
    if bg is not None:
        bgb = ctx.alloc(bg.nbytes)
        bgb.upload(bg.astype(np.float32))
    c4p = np.zeros(n, np.int64) if cam_for_proj is None else cam_for_proj
    X.line_int(ctx, vol, cams, poses_f32(poses), c4p, w.step, buf, kernel=kernel, store=store, bg=bgb,
               default_bg=default_bg, geom=geom)
    out = buf.download(np.float32).reshape(n, rows, cols)
    buf.free()
    if bgb:
        bgb.free()
    vol.close()
    return out


@pytest.fixture(scope="module")
def c1s():
    from xreg_amd import synth

    return synth.config("c1", n_poses=6, small=True)


@pytest.mark.parametrize("layout", [0, 1, 3, 4])
def test_drr_fp32_matches_cpu_reference(ctx, oracle, c1s, layout):
    print(f"case: c1-small layout={layout}", flush=True)
    poses = np.concatenate([c1s.gt[None], c1s.poses])
    ref, s_ref = oracle.drr_f32(c1s.vol, c1s.idx_to_phys, c1s.cams, poses_f32(poses), step=c1s.step)
    got = _gpu_drr(ctx, c1s, poses, layout=layout)
    err = drr_rel_err(got, ref)
    ref64, _ = oracle.drr_f64(c1s.vol, c1s.idx_to_phys, c1s.cams, poses_f32(poses).astype(np.float64),
                              step=c1s.step)
    print(f"max rel err vs FP32 CPU ref {err.max():.3e}; vs FP64 oracle {drr_rel_err(got, ref64).max():.3e}")
    assert err.max() <= DRR_REL_TOL


@pytest.mark.parametrize("name", ["c2", "c3", "c4"])
def test_drr_fp32_matches_cpu_reference_other_configs(ctx, oracle, name):
    """More rays (other geometries) through the FP32 set-up: no step-count flips anywhere."""
    from xreg_amd import synth

    w = synth.config(name, n_poses=8, small=True)
    cams = w.cams[:1]
    print(f"case: {name}-small 8 poses", flush=True)
    ref, s_ref = oracle.drr_f32(w.vol, w.idx_to_phys, cams, poses_f32(w.poses), step=w.step)
    got = _gpu_drr(ctx, w, w.poses, cams=cams)
    err = drr_rel_err(got, ref)
    k = np.unravel_index(np.argmax(err), err.shape)
    print(f"max rel err {err.max():.3e} at {k}: gpu {got[k]!r} cpu {ref[k]!r}")
    assert err.max() <= DRR_REL_TOL


@pytest.mark.parametrize("name,geom", [("c1", 0), ("c2", 0), ("c2", 1), ("c4", 1)])
def test_empty_space_skipping_is_exact(ctx, name, geom, monkeypatch):
    """DRRs with the empty-space map equal DRRs without it bit for bit, and the map removes
    samples (the phantom sits in air)."""
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    w = synth.config(name, n_poses=6, small=True)
    print(f"case: skipping {name}-small geom={geom}", flush=True)
    poses = np.concatenate([w.gt[None], w.poses])
    with_skip = _gpu_drr(ctx, w, poses, geom=geom, cams=w.cams[:1])
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    total, evaluated = X.line_int_sample_count(ctx, vol, w.cams[:1], poses_f32(poses),
                                               np.zeros(poses.shape[0], np.int64), w.step, geom=geom,
                                               evaluated=True)
    vol.close()
    monkeypatch.setenv("XRH_SKIP_EMPTY", "0")
    without = _gpu_drr(ctx, w, poses, geom=geom, cams=w.cams[:1])
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    t2, e2 = X.line_int_sample_count(ctx, vol, w.cams[:1], poses_f32(poses), np.zeros(poses.shape[0], np.int64),
                                     w.step, geom=geom, evaluated=True)
    vol.close()
    print(f"samples {total}, evaluated with skipping {evaluated} ({evaluated / total:.3f})")
    assert np.array_equal(with_skip, without)
    assert t2 == total and e2 == total
    assert evaluated <= total


@pytest.mark.parametrize("geom", [0, 1])
def test_empty_space_skipping_sparse_volume(ctx, geom, monkeypatch):
    """A small dense ball in a large air volume: most samples are skipped and the DRR is
    unchanged bit for bit."""
    from xreg_amd import _xrh as X
    from xreg_amd import geometry as G
    from xreg_amd import synth

    print(f"case: skipping sparse ball geom={geom}", flush=True)
    n = 96
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float32)
    c = (n - 1) / 2.0
    rng = np.random.default_rng(4)
    vol = np.where((x - c) ** 2 + (y - c - 6) ** 2 + (z - c + 4) ** 2 <= 12.0 ** 2,
                   0.02 + 0.01 * rng.random((n, n, n), dtype=np.float32), 0.0).astype(np.float32)
    sp = 1.5
    i2p = G.idx_to_phys(sp)
    center = G.vol_center_phys((n, n, n), sp)
    gt = G.gt_pose(center, 510.0)
    poses = np.concatenate([gt[None], synth.random_poses(gt, center, 5, 15.0, 10.0, seed=9)])
    cam = G.camera(1020.0, 64, 64, 3.0, 3.0)
    w = synth.Workload(name="ball", vol=vol, spacing=sp, cams=[cam], step=0.75, gt=gt, poses=poses,
                       center=center, metric=0)
    with_skip = _gpu_drr(ctx, w, poses, geom=geom)
    v = X.Volume(ctx, vol, i2p)
    total, evaluated = X.line_int_sample_count(ctx, v, [cam], poses_f32(poses), np.zeros(len(poses), np.int64),
                                               w.step, geom=geom, evaluated=True)
    v.close()
    monkeypatch.setenv("XRH_SKIP_EMPTY", "0")
    without = _gpu_drr(ctx, w, poses, geom=geom)
    print(f"samples {total}, evaluated with skipping {evaluated} ({evaluated / total:.3f})")
    assert with_skip.max() > 0
    assert np.array_equal(with_skip, without)
    assert evaluated < 0.5 * total


@pytest.mark.parametrize("geom", [0, 1])
def test_skipping_with_gaps_and_partial_tiles(ctx, oracle, geom, monkeypatch):
    """Every lane of a wave marches from the wave's earliest first sample. Rays here cross two
    balls with air between them, neighbouring rays find their first sample far apart (one ball
    sits off the axis, so a wave holds rays that miss it and rays that graze it), and the
    70x45 detector leaves partial tiles with idle lanes: the DRRs with the empty-space map
    equal those without it bit for bit, in both sample loops (cell reuse or two samples per
    iteration), and match the oracle."""
    from xreg_amd import _xrh as X
    from xreg_amd import geometry as G
    from xreg_amd import synth

    print(f"case: skipping gaps + partial tiles geom={geom}", flush=True)
    n = 80
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float32)
    rng = np.random.default_rng(12)
    balls = ((40.0, 18.0, 40.0, 9.0), (40.0, 62.0, 40.0, 9.0), (58.0, 40.0, 30.0, 6.0))  # (x, y, z, r)
    inside = np.zeros((n, n, n), bool)
    for bx, by, bz, br in balls:
        inside |= (x - bx) ** 2 + (y - by) ** 2 + (z - bz) ** 2 <= br * br
    vol = np.where(inside, 0.02 + 0.01 * rng.random((n, n, n), dtype=np.float32), 0.0).astype(np.float32)
    sp = 1.5
    center = G.vol_center_phys((n, n, n), sp)
    gt = G.gt_pose(center, 510.0)
    # float32 poses throughout: the plain entry takes float poses, the batch takes doubles and
    # its FP64 ray set-up would otherwise see the unrounded values
    poses = poses_f32(np.concatenate([gt[None], synth.random_poses(gt, center, 5, 12.0, 8.0, seed=3)]))
    poses = poses.reshape(-1, 4, 4).astype(np.float64)
    cam = G.camera(1020.0, 45, 70, 2.5, 2.5)
    w = synth.Workload(name="gaps", vol=vol, spacing=sp, cams=[cam], step=0.75, gt=gt, poses=poses,
                       center=center, metric=0)
    i2p = G.idx_to_phys(sp)
    npx = cam.num_det_rows * cam.num_det_cols

    def batched(reuse):
        v = X.Volume(ctx, vol, i2p)
        fx = X.Fixed(ctx, np.ones((cam.num_det_rows, cam.num_det_cols), np.float32), 0)
        b = X.Batch(ctx, [X.make_view(v, fx, cam, w.step)], max_items=len(poses), geom=geom)
        b.set_drr_variant(cell_reuse=reuse)
        buf = ctx.alloc(4 * len(poses) * npx)
        b.eval(np.zeros(len(poses), np.int32), poses, buf)
        out = buf.download(np.float32).reshape(len(poses), cam.num_det_rows, cam.num_det_cols)
        buf.free()
        b.close()
        fx.close()
        v.close()
        return out

    plain = _gpu_drr(ctx, w, poses, geom=geom)
    reuse, pairs = batched(1), batched(0)
    monkeypatch.setenv("XRH_SKIP_EMPTY", "0")
    plain_ns = _gpu_drr(ctx, w, poses, geom=geom)
    reuse_ns = batched(1)
    assert plain.max() > 0
    for got in (plain, reuse, pairs, reuse_ns):
        assert np.array_equal(got, plain_ns)
    if geom == 0:
        ref, _ = oracle.drr_f32(vol, i2p, [cam], poses_f32(poses), step=w.step)
        err = drr_rel_err(plain, ref.reshape(plain.shape))
        print(f"  max rel err vs oracle {err.max():.2e}", flush=True)
        assert err.max() <= DRR_REL_TOL


def test_drr_sample_count_matches_oracle(ctx, oracle, c1s):
    from xreg_amd import _xrh as X

    poses = poses_f32(np.concatenate([c1s.gt[None], c1s.poses]))
    _, s_ref = oracle.drr_f32(c1s.vol, c1s.idx_to_phys, c1s.cams, poses, step=c1s.step)
    vol = X.Volume(ctx, c1s.vol, c1s.idx_to_phys)
    s = X.line_int_sample_count(ctx, vol, c1s.cams, poses, np.zeros(poses.shape[0], np.int64), c1s.step)
    vol.close()
    assert s == s_ref


@pytest.mark.parametrize("name", ["c1", "c2", "c4"])
def test_drr_fp64_geometry_flips_attributed(ctx, oracle, name):
    """FP64 ray set-up (the batched path's default) against the FP32 CPU restatement: every pixel
    off by more than 1e-4 is a step-count flip (the device and the restatement disagree on the
    ray's n) or a pixel where the restatement's own float arithmetic is the larger error (the
    device is at least as close to the FP64 oracle). xrh_line_int with FP64 geometry casts the
    same rays as the batch (identical DRRs), and its step counts come from the batch."""
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    from tests.helpers import attribute_drr_errors, describe_pixels

    w = synth.config(name, n_poses=12, small=True)
    cam = w.cams[0]
    P = cam.num_det_rows * cam.num_det_cols
    poses = poses_f32(np.concatenate([w.gt[None], w.poses])).reshape(-1, 4, 4).astype(np.float64)
    n = len(poses)
    print(f"case: fp64 geometry attribution {name}-small, {n} poses", flush=True)
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    fixed = np.ones((cam.num_det_rows, cam.num_det_cols), np.float32)
    fx = X.Fixed(ctx, fixed, X.XRH_SIM_NCC)
    batch = X.Batch(ctx, [X.make_view(vol, fx, cam, w.step)], max_items=n, geom=1)
    buf = ctx.alloc(4 * P * n)
    iv = np.zeros(n, np.int32)
    batch.eval(iv, poses, drr=buf)
    got = buf.download(np.float32).reshape(n, cam.num_det_rows, cam.num_det_cols)
    batch.ray_steps(iv, poses, buf)
    steps_dev = buf.download(np.int32).reshape(got.shape)
    X.line_int(ctx, vol, [cam], poses_f32(poses), np.zeros(n, np.int64), w.step, buf, geom=1)
    plain = buf.download(np.float32).reshape(got.shape)
    buf.free()
    batch.close()
    fx.close()
    vol.close()
    assert np.array_equal(plain, got)
    pf = poses_f32(poses)
    ref32, _ = oracle.drr_f32(w.vol, w.idx_to_phys, [cam], pf, step=w.step)
    ref64, _ = oracle.drr_f64(w.vol, w.idx_to_phys, [cam], pf.astype(np.float64), step=w.step)
    steps_cpu = oracle.ray_steps(w.vol.shape, w.idx_to_phys, [cam], pf, step=w.step)
    a = attribute_drr_errors(got, ref32, ref64, steps_dev, steps_cpu)
    e64 = drr_rel_err(got, ref64)
    print(f"{a['over_tol']} of {a['pixels']} pixels over 1e-4: {a['flips']} flips, {a['ref_drift']} restatement "
          f"drift, {a['unexplained']} unexplained; max outside flips {a['max_err_not_flip']:.3e}; "
          f"vs FP64 oracle max rel {e64.max():.3e}, median {np.median(e64):.3e}")
    print(f"drift pixels where the device is within 1e-4 (or 1e-6 of the image maximum) of FP64 but not nearer "
          f"than the restatement: {a['ref_drift_within_tol_only']}")
    if a["unexplained"]:
        print("unexplained: " + describe_pixels(a["unexplained_mask"], got, ref32, ref64))
    assert a["unexplained"] == 0
    assert a["flips"] <= 1e-3 * a["pixels"]
    assert np.median(e64) < 1e-6


@pytest.mark.parametrize("storage,layout", [(1, 2), (2, 2), (1, 0)])
def test_drr_half_storage_error_vs_fp64(ctx, oracle, c1s, storage, layout):
    poses = np.concatenate([c1s.gt[None], c1s.poses[:2]])
    ref64, _ = oracle.drr_f64(c1s.vol, c1s.idx_to_phys, c1s.cams, poses_f32(poses).astype(np.float64),
                              step=c1s.step)
    got = _gpu_drr(ctx, c1s, poses, storage=storage, layout=layout)
    # FP16 rounds each voxel by up to 2^-11 relative, BF16 by 2^-8: the storage error of a
    # line integral is bounded by that fraction of the integral of |mu| along the ray, i.e.
    # of the image maximum (grazing pixels inherit it undiluted, hence max-normalised)
    err = np.abs(got - ref64) / ref64.max()
    bound = 2.0 ** -11 if storage == 1 else 2.0 ** -8
    print(f"storage={storage}: max |err| / max(DRR) vs FP64 {err.max():.3e} (bound {bound:.2e})")
    assert err.max() < bound


def test_drr_max_kernel_accum_and_bg(ctx, oracle, c1s):
    poses = np.concatenate([c1s.gt[None], c1s.poses[:2]])
    n = poses.shape[0]
    rows, cols = c1s.cams[0].num_det_rows, c1s.cams[0].num_det_cols
    rng = np.random.default_rng(7)
    init = rng.uniform(0, 1, size=(n, rows, cols)).astype(np.float32)
    bg = rng.uniform(0, 1, size=(1, rows, cols)).astype(np.float32)
    # MAX projection (MIP), REPLACE with default background 0.25
    ref, _ = oracle.drr_f32(c1s.vol, c1s.idx_to_phys, c1s.cams, poses_f32(poses), step=c1s.step, kernel=1,
                            default_bg=0.25)
    got = _gpu_drr(ctx, c1s, poses, kernel=1, default_bg=0.25)
    assert np.array_equal(got, ref)
    # SUM, ACCUM into existing contents
    ref, _ = oracle.drr_f32(c1s.vol, c1s.idx_to_phys, c1s.cams, poses_f32(poses), step=c1s.step, store=1,
                            out=init.copy())
    got = _gpu_drr(ctx, c1s, poses, store=1, init=init)
    assert drr_rel_err(got, ref).max() <= DRR_REL_TOL
    # SUM, REPLACE with a background image per camera
    ref, _ = oracle.drr_f32(c1s.vol, c1s.idx_to_phys, c1s.cams, poses_f32(poses), step=c1s.step, bg=bg)
    got = _gpu_drr(ctx, c1s, poses, bg=bg)
    assert drr_rel_err(got, ref).max() <= DRR_REL_TOL
    # background images with ACCUM: the backgrounds replace the previous contents first
    # (xregRayCastBaseCPU.cpp:134-143 runs before the store switch), so the result equals REPLACE
    ref_acc, _ = oracle.drr_f32(c1s.vol, c1s.idx_to_phys, c1s.cams, poses_f32(poses), step=c1s.step, store=1,
                                bg=bg, out=init.copy())
    assert np.array_equal(ref_acc, ref)
    got_acc = _gpu_drr(ctx, c1s, poses, store=1, init=init, bg=bg)
    assert np.array_equal(got_acc, got)
    # MAX with background images: max(bg, line maximum); a sample may differ from ITK's by an
    # ulp (the trilinear layout's arithmetic), so the DRR bar applies, and the background wins
    # exactly where it is larger
    ref, _ = oracle.drr_f32(c1s.vol, c1s.idx_to_phys, c1s.cams, poses_f32(poses), step=c1s.step, kernel=1, bg=bg)
    got = _gpu_drr(ctx, c1s, poses, kernel=1, bg=bg)
    assert drr_rel_err(got, ref).max() <= DRR_REL_TOL
    assert np.all(got >= bg)
    assert np.count_nonzero(got == bg) > 0 and np.count_nonzero(got > bg) > 0


def test_drr_multiview_extrinsics(ctx, oracle):
    from xreg_amd import synth

    w = synth.config("c3", n_poses=3, small=True)
    poses = np.concatenate([w.gt[None], w.poses])
    n = poses.shape[0]
    # view-major projections: proj = cam * n + pose (distribute_xforms_among_cam_models)
    allp = np.concatenate([poses] * 3)
    c4p = np.repeat(np.arange(3, dtype=np.int64), n)
    ref, _ = oracle.drr_f32(w.vol, w.idx_to_phys, w.cams, poses_f32(allp), c4p, step=w.step)
    got = _gpu_drr(ctx, w, allp, cam_for_proj=c4p)
    assert drr_rel_err(got, ref).max() <= DRR_REL_TOL
    # the three views differ
    assert not np.allclose(got[0], got[n])


@pytest.mark.parametrize("frame", [0, 2])
def test_drr_other_camera_frames(ctx, oracle, c1s, frame):
    """Detector at z = +f (POS_Z) and origin on the detector: the pose is turned half a turn about
    x so the volume sits between source and detector."""
    from xreg_amd import geometry as G

    c0 = c1s.cams[0]
    rows, cols = c0.num_det_rows, c0.num_det_cols
    K = G.naive_intrins(1020.0, rows, cols, c0.det_row_spacing, c0.det_col_spacing)
    K[0, 0] *= -1
    K[1, 1] *= -1
    cam = G.camera_ke(K, np.eye(4), rows, cols, c0.det_row_spacing, c0.det_col_spacing, coord_frame=frame)
    flip = G.rigid(G.rot_x(np.pi))
    poses = np.stack([p @ flip for p in np.concatenate([c1s.gt[None], c1s.poses[:2]])])
    print(f"case: camera frame {frame}, 3 poses", flush=True)
    ref, _ = oracle.drr_f32(c1s.vol, c1s.idx_to_phys, [cam], poses_f32(poses), step=c1s.step)
    got = _gpu_drr(ctx, c1s, poses, cams=[cam])
    assert ref.max() > 0
    assert drr_rel_err(got, ref).max() <= DRR_REL_TOL


def test_hu_volume_matches_host_conversion(ctx, oracle):
    """A volume created from Hounsfield units (converted on the device) casts the same DRR,
    bit for bit, as one created from the oracle's host conversion."""
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    print("case: HU volume c1-small", flush=True)
    w = synth.config("c1", n_poses=3, small=True)
    hu = (w.vol / 0.02683 * 1000.0 - 1000.0).astype(np.float32)  # attenuation -> HU
    att = oracle.hu_to_lin_att(hu)
    poses = poses_f32(w.poses)
    c4p = np.zeros(len(poses), np.int64)
    rows, cols = w.cams[0].num_det_rows, w.cams[0].num_det_cols
    out = []
    for kw in ({"hu": {}}, {}):
        vol = X.Volume(ctx, hu if kw else att, w.idx_to_phys, **kw)
        buf = ctx.alloc(4 * len(poses) * rows * cols)
        X.line_int(ctx, vol, w.cams, poses, c4p, w.step, buf)
        out.append(buf.download(np.float32))
        buf.free()
        vol.close()
    assert out[0].max() > 0
    assert np.array_equal(out[0], out[1])


def test_ct_file_to_drr(ctx, oracle, tmp_path):
    """A Hounsfield-unit CT written as NIfTI, read back with its frame and converted on the
    device, casts the oracle's DRR of the host-converted volume."""
    from xreg_amd import _xrh as X
    from xreg_amd import synth, volio

    print("case: NIfTI HU volume c1-small", flush=True)
    w = synth.config("c1", n_poses=2, small=True)
    hu = (w.vol / 0.02683 * 1000.0 - 1000.0).astype(np.float32)
    i2p = w.idx_to_phys.astype(np.float64)
    volio.write_nifti(tmp_path / "ct.nii.gz", volio.Volume3D(hu, np.diag(i2p)[:3], i2p[:3, 3], np.eye(3)))
    vol, v3 = volio.load_ct(ctx, tmp_path / "ct.nii.gz", hu={})
    assert np.array_equal(v3.idx_to_phys(), w.idx_to_phys.astype(np.float32))
    poses = poses_f32(w.poses)
    rows, cols = w.cams[0].num_det_rows, w.cams[0].num_det_cols
    buf = ctx.alloc(4 * len(poses) * rows * cols)
    X.line_int(ctx, vol, w.cams, poses, np.zeros(len(poses), np.int64), w.step, buf)
    got = buf.download(np.float32).reshape(len(poses), rows, cols)
    buf.free()
    vol.close()
    ref, _ = oracle.drr_f32(oracle.hu_to_lin_att(hu), w.idx_to_phys, w.cams, poses, step=w.step)
    assert ref.max() > 0
    assert drr_rel_err(got, ref).max() <= DRR_REL_TOL


def test_fp32_layouts_agree(ctx, c1s):
    """The coefficient layout evaluates the same trilinear interpolant as the corner layouts
    (rounding aside): DRRs agree to 2e-6 of the image maximum."""
    print("case: layouts 0/1/3 agree c1-small", flush=True)
    poses = np.concatenate([c1s.gt[None], c1s.poses[:3]])
    outs = [_gpu_drr(ctx, c1s, poses, layout=l) for l in (0, 1, 3, 4)]
    m = float(outs[0].max())
    for o in outs[1:]:
        assert np.abs(o.astype(np.float64) - outs[0]).max() <= 2e-6 * m


@pytest.mark.parametrize("layout", [3, 4, 5])
@pytest.mark.parametrize("name", ["c1", "c2", "c4"])
@pytest.mark.parametrize("geom", [0, 1])
@pytest.mark.parametrize("kernel,store", [(0, 0), (1, 0), (0, 1)])
def test_cell_reuse_is_bit_identical(ctx, name, geom, kernel, store, layout, monkeypatch):
    """Sampling that keeps the previous sample's cell coefficients when the cell is unchanged
    (chosen for launches that fill the GPU many times over) equals the kernel that loads every
    sample bit for bit, for SUM and MAX, REPLACE and ACCUM, both ray set-ups, with skipping."""
    from xreg_amd import synth

    print(f"case: cell reuse vs per-sample loads {name}-small geom={geom} kernel={kernel} store={store} "
          f"layout={layout}",
          flush=True)
    w = synth.config(name, n_poses=8, small=True)
    poses = np.concatenate([w.gt[None], w.poses])
    init = np.random.default_rng(1).random((len(poses), w.cams[0].num_det_rows, w.cams[0].num_det_cols))
    kw = dict(geom=geom, kernel=kernel, store=store, cams=w.cams[:1], init=init if store else None, layout=layout)
    monkeypatch.setenv("XRH_CELL_REUSE", "1")
    a = _gpu_drr(ctx, w, poses, **kw)
    monkeypatch.setenv("XRH_CELL_REUSE", "0")
    b = _gpu_drr(ctx, w, poses, **kw)
    assert np.array_equal(a, b)


def test_drr_edge_cases(ctx, oracle):
    """Volume out of view (with and without a default background), volumes of one or two voxels
    per axis, a step longer than every chord (one sample per ray), 1x1 and 3x17 detectors."""
    from xreg_amd import _xrh as X
    from xreg_amd import geometry as G
    from xreg_amd import synth

    w = synth.config("c1", n_poses=2, small=True)

    def both(vol, i2p, cams, poses, step, **kw):
        ref, s_ref = oracle.drr_f32(vol, i2p, cams, poses_f32(poses), step=step, **kw)
        rows, cols = cams[0].num_det_rows, cams[0].num_det_cols
        v = X.Volume(ctx, vol, i2p)
        buf = ctx.alloc(4 * len(poses) * rows * cols)
        X.line_int(ctx, v, cams, poses_f32(poses), np.zeros(len(poses), np.int64), step, buf, **kw)
        got = buf.download(np.float32).reshape(ref.shape)
        buf.free()
        v.close()
        return got, ref

    print("case: drr edge cases", flush=True)
    far = G.rigid(None, (2000.0, 0.0, 0.0)) @ w.gt
    got, ref = both(w.vol, w.idx_to_phys, w.cams, far[None], w.step)
    assert np.array_equal(got, ref) and not got.any()
    got, ref = both(w.vol, w.idx_to_phys, w.cams, far[None], w.step, default_bg=0.25)
    assert np.array_equal(got, ref) and np.all(got == 0.25)
    for shape in [(1, 1, 1), (1, 2, 3), (2, 2, 2), (3, 1, 4)]:
        vol = np.random.default_rng(0).uniform(0.1, 1.0, shape).astype(np.float32)
        i2p = G.idx_to_phys(8.0)
        gt = G.gt_pose(G.vol_center_phys(shape[::-1], 8.0), 510.0)
        got, ref = both(vol, i2p, w.cams, gt[None], 1.0)
        assert drr_rel_err(got, ref).max() <= DRR_REL_TOL, shape
    const = np.full(w.vol.shape, 0.5, np.float32)
    got, ref = both(const, w.idx_to_phys, w.cams, w.gt[None], 1000.0)
    assert ref.max() == np.float32(500.0) and np.array_equal(got, ref)
    got, ref = both(w.vol, w.idx_to_phys, [G.camera(1020.0, 1, 1, 0.776, 0.776)], w.gt[None], w.step)
    assert ref[0, 0, 0] > 0 and drr_rel_err(got, ref).max() <= DRR_REL_TOL
    got, ref = both(w.vol, w.idx_to_phys, [G.camera(1020.0, 3, 17, 4.0, 2.0)], w.poses, w.step)
    assert drr_rel_err(got, ref).max() <= DRR_REL_TOL


@pytest.mark.parametrize("geom", [0, 1])
def test_brick_layout_equals_face_layout(ctx, oracle, geom):
    """The bricked x-face layout holds the same coefficients as the x-face layout, so its DRRs
    are the same bit for bit (both loop forms), at an odd volume size whose bricks overhang
    the volume on every axis."""
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    w = synth.config("c2", n_poses=5, small=True)
    vol_zyx = np.ascontiguousarray(w.vol[:47, :61, :57])
    w2 = synth.Workload(name="odd", vol=vol_zyx, spacing=w.spacing, cams=w.cams, step=w.step, gt=w.gt,
                        poses=w.poses, center=w.center)
    poses = np.concatenate([w.gt[None], w.poses])
    print(f"case: brick vs face layout, volume {vol_zyx.shape[::-1]}, geom {geom}", flush=True)
    face = _gpu_drr(ctx, w2, poses, layout=X.LAYOUT_FACE, geom=geom)
    brick = _gpu_drr(ctx, w2, poses, layout=X.LAYOUT_BRICK, geom=geom)
    assert face.max() > 0
    assert np.array_equal(face, brick)
    if geom == 0:
        ref, _ = oracle.drr_f32(w2.vol, w2.idx_to_phys, w2.cams, poses_f32(poses), step=w2.step)
        assert drr_rel_err(brick, ref).max() <= DRR_REL_TOL
