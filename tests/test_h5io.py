"""HDF5 projection data, camera, parameterisation and debug-record files (CPU only).

Pinned three ways: round trips through this package's writer and reader; the layout of a
written file as printed by the HDF5 distribution's own h5dump (dataset names, on-disk types,
shapes, chunking and deflate level of the reference's writers); and reading a file that the
HDF5 distribution's h5import built without this package."""
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from xreg_amd import geometry as G
from xreg_amd import optvars as OV

h5io = pytest.importorskip("xreg_amd.h5io")
if not h5io.available():
    pytest.skip("HDF5 C library not available", allow_module_level=True)


def _tool(name):
    for d in [os.path.dirname(os.path.realpath(h5io.lib().h._name)) + "/../bin", "/opt/conda/bin"]:
        p = Path(d) / name
        if p.exists():
            return str(p)
    return shutil.which(name)


def _cam_bytes(c):
    return bytes(memoryview(c))


def _projs(dtype, seed=0):
    rng = np.random.default_rng(seed)
    E = G.rigid(G.rot_y(np.deg2rad(30.0)), (5.0, -3.0, 2.0))
    cams = [G.camera(1020.0, 48, 40, 0.776, 0.776),
            G.camera(1020.0, 48, 40, 0.776, 0.776, extrins=E),
            G.camera_ke(G.naive_intrins(980.0, 48, 40, 0.5, 0.6), E, 48, 40, 0.5, 0.6, coord_frame=G.FRAME_ON_DETECTOR)]
    out = []
    for i, c in enumerate(cams):
        if np.dtype(dtype) == np.float32:
            px = rng.standard_normal((48, 40)).astype(np.float32)
        else:
            px = rng.integers(0, np.iinfo(dtype).max, (48, 40)).astype(dtype)
        img = h5io.Image2D(px, spacing=np.array([c.det_col_spacing, c.det_row_spacing], np.float32),
                           origin=np.array([1.5, -2.0], np.float32))
        pd = h5io.ProjData(cam=c, img=img if i != 1 or dtype != np.float32 else img)
        if i == 0:
            pd.landmarks = {"FH-l": np.array([10.5, 20.25], np.float32), "GSN-r": np.array([3.0, 4.0], np.float32)}
            pd.rot_to_pat_up = 90
            pd.det_spacings_from_orig_meta = True
            pd.orig_dcm_meta = {"patient-id": "anon", "rows": np.uint64(48)}
        out.append(pd)
    return out


@pytest.mark.parametrize("dtype", [np.float32, np.uint16, np.uint8])
def test_proj_data_round_trip(tmp_path, dtype):
    projs = _projs(dtype)
    path = tmp_path / "pd.h5"
    h5io.write_proj_data(path, projs)
    got = h5io.read_proj_data(path)
    assert len(got) == len(projs)
    for a, b in zip(projs, got):
        # reading rebuilds the camera with CameraModel::setup(K, E, ...), as ReadCamModelH5 does
        # (a naive camera's identity extrinsic inverse then carries -0 translations)
        want = G.camera_ke(a.cam.K, a.cam.E, a.cam.num_det_rows, a.cam.num_det_cols, a.cam.det_row_spacing,
                           a.cam.det_col_spacing, coord_frame=a.cam.coord_frame)
        assert _cam_bytes(want) == _cam_bytes(b.cam)
        for f in ("intrins_inv", "extrins_inv", "pinhole"):
            assert np.array_equal(getattr(a.cam, f)[:], getattr(b.cam, f)[:])
        assert np.array_equal(a.cam.K, b.cam.K) and np.array_equal(a.cam.E, b.cam.E)
        assert b.img.pixels.dtype == np.dtype(dtype)
        assert np.array_equal(a.img.pixels, b.img.pixels)
        assert np.array_equal(a.img.spacing, b.img.spacing) and np.array_equal(a.img.origin, b.img.origin)
        assert np.array_equal(b.img.direction, np.eye(2))
        assert set(a.landmarks) == set(b.landmarks)
        for k in a.landmarks:
            assert np.array_equal(a.landmarks[k], b.landmarks[k])
        assert a.rot_to_pat_up == b.rot_to_pat_up
        assert a.det_spacings_from_orig_meta == b.det_spacings_from_orig_meta
    assert got[0].orig_dcm_meta["patient-id"] == "anon"
    assert got[0].orig_dcm_meta["rows"].shape == () and int(got[0].orig_dcm_meta["rows"]) == 48
    assert h5io.proj_data_scalar_type(path) == h5io.SCALAR_TYPES[np.dtype(dtype)]
    # typed reads convert on read (ReadProjDataH5F32 of a uint16 file)
    f32 = h5io.read_proj_data(path, dtype=np.float32)
    assert np.array_equal(f32[2].img.pixels, projs[2].img.pixels.astype(np.float32))
    # cameras only / one image
    assert all(p.img is None for p in h5io.read_proj_data(path, read_pixels=False))
    assert np.array_equal(h5io.read_single_img(path, 1).pixels, projs[1].img.pixels)
    with pytest.raises(h5io.H5Error):
        h5io.read_single_img(path, 3)


def test_copy_lands_and_deferred_reader(tmp_path):
    projs = _projs(np.float32, seed=1)
    src, dst = tmp_path / "a.h5", tmp_path / "b.h5"
    h5io.write_proj_data(src, projs, compress=False)
    h5io.copy_proj_data(src, dst)
    h5io.add_lands_to_proj_data(dst, {"ASIS-l": [7.0, 8.0]}, 2)
    h5io.add_lands_to_proj_data(dst, {"new": [1.0, 2.0]}, 0, delete_existing=True)
    got = h5io.read_proj_data(dst)
    assert set(got[0].landmarks) == {"new"}
    assert np.array_equal(got[2].landmarks["ASIS-l"], [7.0, 8.0])
    r = h5io.DeferredProjReader(dst, cache_imgs=True)
    assert r.num_projs_on_disk() == 3 and r.scalar_type_on_disk == "float32"
    a = r.read_proj(1)
    assert np.array_equal(a.pixels, projs[1].img.pixels)
    assert r.read_proj(1) is a


def test_mixed_pixel_types_rejected(tmp_path):
    p = _projs(np.float32)[:1] + _projs(np.uint16)[1:2]
    path = tmp_path / "mixed.h5"
    h5io.write_proj_data(path, p)
    with pytest.raises(h5io.H5Error, match="Inconsistent"):
        h5io.proj_data_scalar_type(path)


def test_written_layout_as_h5dump_prints_it(tmp_path):
    """Dataset names, on-disk types, shapes, chunks and deflate levels of the reference's
    writers, read back by h5dump (an HDF5 tool, independent of this package)."""
    tool = _tool("h5dump")
    if tool is None:
        pytest.skip("h5dump not available")
    projs = _projs(np.uint16)
    big = np.arange(1536 * 1536, dtype=np.float32).reshape(1536, 1536)
    projs[1].img = h5io.Image2D(big)
    path = tmp_path / "pd.h5"
    h5io.write_proj_data(path, projs)
    txt = subprocess.run([tool, "-p", "-H", str(path)], capture_output=True, text=True, check=True).stdout
    flat = " ".join(txt.split())
    assert 'ATTRIBUTE "xreg-type" { DATATYPE H5T_STRING { STRSIZE 9;' in flat
    assert 'DATASET "num-projs" { DATATYPE H5T_STD_U64LE DATASPACE SCALAR' in flat
    for g in ["proj-000", "proj-001", "proj-002"]:
        assert f'GROUP "{g}"' in flat
    # camera datasets (xregH5CamModelIO.cpp:29-58)
    assert 'DATASET "intrinsic" { DATATYPE H5T_IEEE_F32LE DATASPACE SIMPLE { ( 3, 3 ) / ( 3, 3 ) }' in flat
    assert 'DATASET "extrinsic" { DATATYPE H5T_IEEE_F32LE DATASPACE SIMPLE { ( 4, 4 ) / ( 4, 4 ) }' in flat
    assert 'DATASET "num-rows" { DATATYPE H5T_STD_U64LE DATASPACE SCALAR' in flat
    assert 'DATASET "row-spacing" { DATATYPE H5T_IEEE_F32LE DATASPACE SCALAR' in flat
    assert 'DATASET "cam-coord-frame-type" { DATATYPE H5T_STRING { STRSIZE 28;' in flat
    assert 'DATASET "rot-to-pat-up" { DATATYPE H5T_STD_I32LE DATASPACE SCALAR' in flat
    assert 'DATASET "det-spacings-from-orig-meta" { DATATYPE H5T_STD_U8LE DATASPACE SCALAR' in flat
    # image: 48x40 uint16 fits one chunk; 1536^2 float32 (9 MiB) is cut to 170 whole rows
    assert 'DATASET "pixels" { DATATYPE H5T_STD_U16LE DATASPACE SIMPLE { ( 48, 40 ) / ( 48, 40 ) } ' \
           'STORAGE_LAYOUT { CHUNKED ( 48, 40 )' in flat
    assert 'DATASET "pixels" { DATATYPE H5T_IEEE_F32LE DATASPACE SIMPLE { ( 1536, 1536 ) / ( 1536, 1536 ) } ' \
           'STORAGE_LAYOUT { CHUNKED ( 170, 1536 )' in flat
    assert flat.count("COMPRESSION DEFLATE { LEVEL 9 }") == 3
    assert 'DATASET "origin" { DATATYPE H5T_IEEE_F32LE DATASPACE SIMPLE { ( 2, 1 ) / ( 2, 1 ) } STORAGE_LAYOUT { CONTIGUOUS' in flat
    assert 'DATASET "FH-l" { DATATYPE H5T_IEEE_F32LE DATASPACE SIMPLE { ( 2, 1 ) / ( 2, 1 ) }' in flat


def test_reads_file_built_by_h5import(tmp_path):
    """A projection file assembled by h5import (variable-length strings and one-element
    vectors instead of scalars, as other HDF5 writers produce) reads into the same camera as
    CameraModel::setup."""
    tool = _tool("h5import")
    if tool is None:
        pytest.skip("h5import not available")
    K = G.naive_intrins(1020.0, 6, 5, 0.5, 0.5)
    E = G.rigid(G.rot_x(0.3), (1.0, 2.0, 3.0)).astype(np.float32)
    px = np.arange(30, dtype=np.uint16).reshape(6, 5)
    items = [
        ("num-projs", "TEXTIN", [1], "UIN", 64, "1"),
        ("proj-000/cam/num-rows", "TEXTIN", [1], "UIN", 64, "6"),
        ("proj-000/cam/num-cols", "TEXTIN", [1], "UIN", 64, "5"),
        ("proj-000/cam/row-spacing", "TEXTFP", [1], "FP", 32, "0.5"),
        ("proj-000/cam/col-spacing", "TEXTFP", [1], "FP", 32, "0.5"),
        ("proj-000/cam/intrinsic", "TEXTFP", [3, 3], "FP", 32, " ".join(repr(float(x)) for x in K.reshape(-1))),
        ("proj-000/cam/extrinsic", "TEXTFP", [4, 4], "FP", 32, " ".join(repr(float(x)) for x in E.reshape(-1))),
        ("proj-000/img/pixels", "TEXTIN", [6, 5], "UIN", 16, " ".join(str(int(x)) for x in px.reshape(-1))),
        ("proj-000/img/spacing", "TEXTFP", [2, 1], "FP", 32, "0.5 0.5"),
        ("proj-000/img/origin", "TEXTFP", [2, 1], "FP", 32, "0 0"),
        ("proj-000/img/dir-mat", "TEXTFP", [2, 2], "FP", 32, "1 0 0 1"),
    ]
    args = []
    for k, (path, cls, dims, ocls, size, text) in enumerate(items):
        (tmp_path / f"{k}.txt").write_text(text + "\n")
        (tmp_path / f"{k}.cfg").write_text(
            f"PATH {path}\nINPUT-CLASS {cls}\nRANK {len(dims)}\nDIMENSION-SIZES {' '.join(map(str, dims))}\n"
            f"OUTPUT-CLASS {ocls}\nOUTPUT-SIZE {size}\nOUTPUT-BYTE-ORDER LE\n")
        args += [str(tmp_path / f"{k}.txt"), "-c", str(tmp_path / f"{k}.cfg")]
    (tmp_path / "s.txt").write_text("origin-at-focal-pt-det-neg-z\n")
    (tmp_path / "s.cfg").write_text("PATH proj-000/cam/cam-coord-frame-type\nINPUT-CLASS STR\n")
    args += [str(tmp_path / "s.txt"), "-c", str(tmp_path / "s.cfg")]
    out = tmp_path / "imp.h5"
    subprocess.run([tool, *args, "-o", str(out)], check=True, capture_output=True)
    pd = h5io.read_proj_data(out)
    assert len(pd) == 1
    want = G.camera_ke(K, E, 6, 5, 0.5, 0.5)
    assert _cam_bytes(pd[0].cam) == _cam_bytes(want)
    assert pd[0].img.pixels.dtype == np.uint16 and np.array_equal(pd[0].img.pixels, px)
    assert h5io.proj_data_scalar_type(out) == "uint16"


def test_camera_frames_follow_camera_model_setup():
    """Pinhole of CameraModel::setup (xregPerspectiveXform.cpp:300-330): the extrinsic inverse's
    translation for focal-point frames, extrins_inv * (0, 0, f) for the detector frame."""
    E = G.rigid(G.rot_z(0.4) @ G.rot_x(-0.2), (3.0, -4.0, 12.0))
    K = G.naive_intrins(1000.0, 10, 12, 0.4, 0.4)
    cn = G.camera_ke(K, E, 10, 12, 0.4, 0.4, coord_frame=G.FRAME_NEG_Z)
    cd = G.camera_ke(K, E, 10, 12, 0.4, 0.4, coord_frame=G.FRAME_ON_DETECTOR)
    einv = np.linalg.inv(E)
    assert np.allclose(np.array(cn.pinhole[:]), einv[:3, 3], atol=1e-4)
    assert np.allclose(np.array(cd.pinhole[:]), einv[:3, :3] @ [0, 0, cd.focal_len] + einv[:3, 3], atol=1e-3)
    assert cd.coord_frame == G.FRAME_ON_DETECTOR
    with pytest.raises(ValueError):
        G.camera_ke(K, E, 10, 12, 0.4, 0.4, coord_frame=7)


@pytest.mark.parametrize("ov", [OV.SE3OptVarsLieAlg(), OV.SE3OptVarsEuler(), OV.SE3OptVarsEuler(0, 1, 2, 3, 4, 5),
                                OV.SE3OptVarsTransOnly(True, False, True), OV.SE3OptVarsTransOnly(False, True, False),
                                OV.SO3OptVarsLieAlg(), OV.SO3OptVarsEuler(0, 2, 1), OV.SO3OptVarsOnly(2)])
def test_se3_opt_vars_round_trip(tmp_path, ov):
    path = tmp_path / "ov.h5"
    with h5io.create_file(path) as f:
        h5io.write_se3_opt_vars(f, ov)
    with h5io.open_file(path) as f:
        got = h5io.read_se3_opt_vars(f)
        assert int(f.scalar("num-params", np.uint64)) == ov.num_params
    assert type(got) is type(ov) and got.num_params == ov.num_params
    x = np.linspace(-0.3, 0.4, ov.num_params)
    assert np.array_equal(got(x), ov(x))


def test_regi_debug_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    nits = 37
    r = h5io.RegiDebugResults(
        se3_params=OV.SE3OptVarsEuler(), vols_used=[0, 2],
        init_poses=[np.eye(4, dtype=np.float32), G.rigid(G.rot_y(0.1), (1, 2, 3)).astype(np.float32)],
        final_poses=[G.rigid(t=(0.5, 0, 0)).astype(np.float32)] * 2,
        inter_frames=[np.eye(4, dtype=np.float32)] * 2, inter_frames_wrt_vol=[True, False],
        iter_vars=[rng.standard_normal((nits, 6)).astype(np.float32) for _ in range(2)],
        sims=rng.random(nits).astype(np.float32), projs_used=[1, 0],
        static_vols=[1], static_vol_poses=[np.eye(4, dtype=np.float32)], regi_time_secs=2.5)
    path = tmp_path / "dbg.h5"
    h5io.write_regi_debug(path, r)
    got = h5io.read_regi_debug(path)
    assert got.vols_used == [0, 2] and got.projs_used == [1, 0] and got.static_vols == [1]
    assert got.inter_frames_wrt_vol == [True, False] and got.regi_time_secs == 2.5
    for a, b in zip(r.iter_vars + r.init_poses + r.final_poses, got.iter_vars + got.init_poses + got.final_poses):
        assert np.array_equal(a, b)
    assert np.array_equal(got.sims, r.sims)
    with h5io.open_file(path) as f:
        with f.open_group("iter-pose-vars") as g:
            assert g.storage("obj-000") == {"chunk": (37, 6), "deflate": 9}
        assert f.storage("sim-vals") == {"chunk": (37,), "deflate": 9}
        with f.open_group("init-poses") as g:
            assert g.storage("obj-001") == {"chunk": None, "deflate": None}


def test_regi_debug_optimizer_and_metric_aux_groups(tmp_path):
    """The optional "opt-aux" (CMA-ES and exhaustive records) and "sims-aux" (patch metrics)
    groups of the debug record: names, on-disk types and shapes of the reference's writers
    (xregIntensity2D3DRegiCMAES.cpp:437-509, xregIntensity2D3DRegiExhaustive.cpp:397-440,
    xregImgSimMetric2DPatchNCCCPU.cpp:443-485, xregImgSimMetric2DPatchGradNCCCPU.cpp:411-420)."""
    rng = np.random.default_rng(5)
    nits, lam = 3, 4
    cma = {"name": "CMA-ES", "pop_size": lam, "obj_dims": [6, 6],
           "pop_params": [rng.standard_normal((lam, 12)) for _ in range(nits)],
           "sim_vals": [rng.random(lam) for _ in range(nits)],
           "cov_mats": [np.eye(12) * (i + 1) for i in range(nits)], "num_rejects": [0, 2, 1]}
    patches = {"start_rows": np.array([0, 0, 1], np.int32), "start_cols": np.array([0, 1, 0], np.int32),
               "stop_rows": np.array([2, 2, 3], np.int32), "stop_cols": np.array([2, 3, 2], np.int32),
               "weights": np.array([1, 1, 1], np.float32)}
    pn = {"type": "patch-ncc-aux", "calls": 2, "patches": patches, "patch_inds": np.arange(3, dtype=np.uint64)}
    pg = {"type": "patch-grad-ncc-aux", "x": pn, "y": pn}
    r = h5io.RegiDebugResults(
        se3_params=OV.SE3OptVarsEuler(), vols_used=[0, 1], init_poses=[np.eye(4, dtype=np.float32)] * 2,
        final_poses=[np.eye(4, dtype=np.float32)] * 2, inter_frames=[np.eye(4, dtype=np.float32)] * 2,
        inter_frames_wrt_vol=[True, True], iter_vars=[rng.standard_normal((nits, 6)).astype(np.float32)] * 2,
        sims=rng.random(nits).astype(np.float32), opt_aux=cma, sims_aux=[pg, None, pn])
    path = tmp_path / "dbg_aux.h5"
    h5io.write_regi_debug(path, r)
    with h5io.open_file(path) as f:
        with f.open_group("opt-aux") as ag:
            assert ag.read("name", str) == "CMA-ES"
            assert int(ag.scalar("num-iterations", np.uint64)) == nits and int(ag.scalar("pop-size", np.uint64)) == lam
            sv = ag.read("all-sim-vals", np.float32)
            assert sv.shape == (nits, lam) and np.allclose(sv, np.stack(cma["sim_vals"]), atol=1e-7)
            with ag.open_group("all-pop-params") as pp, pp.open_group("iter-001") as it:
                assert np.allclose(it.read("obj-001-pop", np.float32), cma["pop_params"][1][:, 6:], atol=1e-6)
            with ag.open_group("covariances") as cg:
                assert np.array_equal(cg.read("iter-002", np.float32), np.eye(12, dtype=np.float32) * 3)
            assert list(ag.read("all-num-rejects", np.uint64)) == [0, 2, 1]
        with f.open_group("sims-aux") as sa:
            assert "001" not in sa
            with sa.open_group("000") as g0:
                assert g0.read("sim-aux-type", str) == "patch-grad-ncc-aux"
                with g0.open_group("y") as gy, gy.open_group("0001") as c1:
                    assert list(c1.read("stop-cols", np.uint64)) == [2, 3, 2]
                    assert list(c1.read("patch-inds", np.uint64)) == [0, 1, 2]
            with sa.open_group("002") as g2:
                assert g2.read("sim-aux-type", str) == "patch-ncc-aux"
                assert "0001" in g2 and "0002" not in g2
    # the exhaustive record
    ex = {"name": "Exhaustive", "cam_wrt_vols": [np.stack([np.eye(4)] * 3)], "has_min": True, "min_sim_val": 0.25,
          "has_all_sim_vals": True, "min_sim_val_idx": 2, "all_sim_vals": np.array([0.5, 0.75, 0.25], np.float32)}
    r.opt_aux, r.sims_aux = ex, []
    h5io.write_regi_debug(tmp_path / "dbg_ex.h5", r)
    with h5io.open_file(tmp_path / "dbg_ex.h5") as f, f.open_group("opt-aux") as ag:
        assert ag.read("name", str) == "Exhaustive"
        assert float(ag.scalar("min-sim-val", np.float32)) == 0.25 and int(ag.scalar("min-sim-val-idx", np.uint64)) == 2
        with ag.open_group("cam-wrt-vols") as cw, cw.open_group("vol-00") as vg:
            assert np.array_equal(vg.read("000002", np.float32), np.eye(4, dtype=np.float32))
        assert "sims-aux" not in f
    tool = _tool("h5dump")
    if tool is not None:
        txt = " ".join(subprocess.run([tool, "-H", str(path)], capture_output=True, text=True, check=True).stdout.split())
        assert 'DATASET "all-sim-vals" { DATATYPE H5T_IEEE_F32LE DATASPACE SIMPLE { ( 3, 4 ) / ( 3, 4 ) }' in txt
        assert 'DATASET "all-num-rejects" { DATATYPE H5T_STD_U64LE' in txt
        assert 'DATASET "weights" { DATATYPE H5T_IEEE_F32LE' in txt and 'DATASET "start-rows" { DATATYPE H5T_STD_U64LE' in txt
