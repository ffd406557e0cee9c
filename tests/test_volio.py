"""CT volume files (NIfTI-1, MetaImage) in ITK's LPS frame, on the CPU.

Headers are assembled byte by byte here from the NIfTI-1 layout (nifti1.h), independently of
the writer, and the expected frames follow the specification's written-out formulas: for the
qform, RAS = R(quaternion) diag(pixdim) diag(1, 1, qfac) ijk + qoffset; ITK negates the first
two coordinates (RAS -> LPS)."""
import gzip
import struct

import numpy as np
import pytest

from xreg_amd import volio


def _nifti_bytes(shape_xyz, dtype_code, bitpix, data_bytes, pixdim, qform=None, sform=None, slope=0.0, inter=0.0):
    hdr = bytearray(352)
    struct.pack_into("<i", hdr, 0, 348)
    X, Y, Z = shape_xyz
    struct.pack_into("<8h", hdr, 40, 3, X, Y, Z, 1, 1, 1, 1)
    struct.pack_into("<2h", hdr, 70, dtype_code, bitpix)
    struct.pack_into("<8f", hdr, 76, *pixdim)
    struct.pack_into("<3f", hdr, 108, 352.0, slope, inter)
    qcode = 1 if qform is not None else 0
    scode = 1 if sform is not None else 0
    struct.pack_into("<2h", hdr, 252, qcode, scode)
    if qform is not None:
        struct.pack_into("<6f", hdr, 256, *qform)
    if sform is not None:
        struct.pack_into("<12f", hdr, 280, *np.asarray(sform, np.float32).reshape(-1))
    hdr[344:348] = b"n+1\x00"
    return bytes(hdr) + data_bytes


def test_nifti_qform_rotation_about_z(tmp_path):
    """Quaternion (b, c, d) = (0, 0, sin 45): a quarter turn about z, qfac = -1."""
    data = np.arange(4 * 3 * 2, dtype=np.int16).reshape(2, 3, 4)
    s = np.sqrt(0.5)
    raw = _nifti_bytes((4, 3, 2), 4, 16, data.tobytes(), (-1.0, 0.5, 0.6, 0.7, 1, 1, 1, 1),
                       qform=(0.0, 0.0, s, 10.0, 20.0, 30.0))
    p = tmp_path / "q.nii"
    p.write_bytes(raw)
    v = volio.read_nifti(p)
    assert np.array_equal(v.data, data) and v.data.dtype == np.int16
    np.testing.assert_allclose(v.spacing, [0.5, 0.6, 0.7], rtol=1e-6)
    # RAS of index (i, j, k): (-0.6 j + 10, 0.5 i + 20, -0.7 k + 30); LPS negates x and y
    M = v.idx_to_phys().astype(np.float64)
    for ijk in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (3, 2, 1)]:
        i, j, k = ijk
        ras = np.array([-0.6 * j + 10.0, 0.5 * i + 20.0, -0.7 * k + 30.0])
        want = np.array([-ras[0], -ras[1], ras[2]])
        got = M[:3, :3] @ np.array(ijk, float) + M[:3, 3]
        np.testing.assert_allclose(got, want, atol=1e-5)


def test_nifti_sform_only_and_scaling(tmp_path):
    """No qform: the sform rows give spacing (column norms) and direction; scl_slope /
    scl_inter scale the stored integers into float."""
    data = np.array([[[0, 100], [200, 300]]], dtype=np.uint16)  # Z=1, Y=2, X=2
    sform = [[0.0, 0.0, 2.0, -5.0], [1.5, 0.0, 0.0, 7.0], [0.0, -1.0, 0.0, 3.0]]
    raw = _nifti_bytes((2, 2, 1), 512, 16, data.tobytes(), (1.0, 9, 9, 9, 1, 1, 1, 1), sform=sform,
                       slope=0.5, inter=-10.0)
    p = tmp_path / "s.nii.gz"
    with gzip.open(p, "wb") as f:
        f.write(raw)
    v = volio.read_nifti(p)
    assert v.data.dtype == np.float32
    assert np.array_equal(v.data, data.astype(np.float64) * 0.5 - 10.0)
    np.testing.assert_allclose(v.spacing, [1.5, 1.0, 2.0])
    M = v.idx_to_phys().astype(np.float64)
    A = np.array(sform)
    for ijk in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]:
        ras = A[:, :3] @ np.array(ijk, float) + A[:, 3]
        np.testing.assert_allclose(M[:3, :3] @ np.array(ijk, float) + M[:3, 3], [-ras[0], -ras[1], ras[2]], atol=1e-5)


def test_nifti_without_transform_is_diagonal(tmp_path):
    data = np.zeros((2, 2, 2), np.float32)
    raw = _nifti_bytes((2, 2, 2), 16, 32, data.tobytes(), (1.0, 0.8, 0.9, 1.2, 1, 1, 1, 1))
    p = tmp_path / "n.nii"
    p.write_bytes(raw)
    v = volio.read_nifti(p)
    np.testing.assert_allclose(v.direction, np.diag([-1.0, -1.0, 1.0]))
    np.testing.assert_allclose(v.spacing, [0.8, 0.9, 1.2], rtol=1e-6)


@pytest.mark.parametrize("dtype", [np.int16, np.float32, np.uint8])
@pytest.mark.parametrize("ext", [".nii", ".nii.gz", ".mha", ".mha.z"])
def test_round_trip(tmp_path, dtype, ext):
    rng = np.random.default_rng(1)
    data = (rng.integers(0, 200, (5, 4, 3)) if dtype != np.float32 else rng.standard_normal((5, 4, 3))).astype(dtype)
    th = 0.4
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    Rx = np.array([[1, 0, 0], [0, np.cos(0.3), -np.sin(0.3)], [0, np.sin(0.3), np.cos(0.3)]])
    v = volio.Volume3D(data, np.array([0.5, 0.75, 1.25]), np.array([-12.0, 4.5, 100.0]), Rz @ Rx)
    if ext.startswith(".nii"):
        p = tmp_path / f"v{ext}"
        volio.write_nifti(p, v)
    else:
        p = tmp_path / "v.mha"
        volio.write_metaimage(p, v, compress=ext.endswith(".z"))
    got = volio.read_volume(p)
    assert got.data.dtype == np.dtype(dtype) and np.array_equal(got.data, data)
    np.testing.assert_allclose(got.spacing, v.spacing, rtol=1e-6)
    np.testing.assert_allclose(got.origin, v.origin, atol=1e-5)
    np.testing.assert_allclose(got.direction, v.direction, atol=1e-6)
    np.testing.assert_allclose(got.idx_to_phys(), v.idx_to_phys(), atol=1e-5)


def test_metaimage_header_conventions(tmp_path):
    """TransformMatrix lists index-axis directions one after another (axis 0 first); data may
    sit in a separate raw file."""
    data = np.arange(8, dtype=np.int16).reshape(2, 2, 2)
    (tmp_path / "v.raw").write_bytes(data.tobytes())
    (tmp_path / "v.mhd").write_text("ObjectType = Image\nNDims = 3\nBinaryData = True\n"
                                    "BinaryDataByteOrderMSB = False\nTransformMatrix = 0 1 0 -1 0 0 0 0 1\n"
                                    "Offset = 1 2 3\nElementSpacing = 2 3 4\nDimSize = 2 2 2\n"
                                    "ElementType = MET_SHORT\nElementDataFile = v.raw\n")
    v = volio.read_volume(tmp_path / "v.mhd")
    assert np.array_equal(v.data, data)
    M = v.idx_to_phys().astype(np.float64)
    # index axis 0 points along +y (2 mm per voxel), axis 1 along -x (3 mm)
    np.testing.assert_allclose(M[:3, :3] @ [1, 0, 0] + M[:3, 3], [1, 4, 3])
    np.testing.assert_allclose(M[:3, :3] @ [0, 1, 0] + M[:3, 3], [-2, 2, 3])
    np.testing.assert_allclose(M[:3, :3] @ [0, 0, 1] + M[:3, 3], [1, 2, 7])


def test_rejections(tmp_path):
    p = tmp_path / "bad.nii"
    p.write_bytes(b"\x00" * 400)
    with pytest.raises(ValueError):
        volio.read_nifti(p)
    with pytest.raises(ValueError):
        volio.read_volume(tmp_path / "x.nrrd")
