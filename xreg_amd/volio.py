"""CT volumes from NIfTI-1 (.nii, .nii.gz) and MetaImage (.mha, .mhd) files, in ITK's
physical frame.

The reference reads volumes with ITK (ReadITKImageFromDisk, lib/itk/xregITKIOUtils.h:113-122;
the pelvis app reads the CT that way and converts it with HUToLinAtt before registering,
apps/hip_surgery/pelvis_single_view_regi_2d_3d/..._main.cpp:143-147). ITK is not part of
this image, so the two formats the reference's data uses are read here from their published
specifications:

* NIfTI-1 (NIfTI Data Format Working Group, nifti1.h): 348-byte header, voxel data at
  vox_offset, optional scl_slope / scl_inter scaling. The index -> RAS transform is the qform
  (quaternion + qoffset + pixdim, qfac = pixdim[0]) when qform_code > 0, else the sform rows
  when sform_code > 0, else diag(pixdim). ITK reports LPS: the first two rows of the RAS
  transform are negated. Spacing is |pixdim[1..3]|; the direction is the transform's
  normalised columns.
* MetaImage (MetaIO): "Key = Value" header; TransformMatrix lists the axis direction
  vectors one after another (direction column i = values 3i .. 3i+2), Offset is the origin,
  ElementSpacing the spacing; data LOCAL (after the header) or in a named file, optionally
  zlib-compressed.

A file whose qform and sform disagree is read with the qform; which one ITK 5.1 picks then is
not pinned here. Everything runs on the host.
"""
from __future__ import annotations

import gzip
import os
import struct
import zlib
from dataclasses import dataclass, field

import numpy as np

_NIFTI_DTYPES = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64, 256: np.int8,
                 512: np.uint16, 768: np.uint32, 1024: np.int64, 1280: np.uint64}
_NIFTI_CODES = {np.dtype(v).str[1:]: k for k, v in _NIFTI_DTYPES.items()}

_MET_TYPES = {"MET_UCHAR": np.uint8, "MET_CHAR": np.int8, "MET_USHORT": np.uint16, "MET_SHORT": np.int16,
              "MET_UINT": np.uint32, "MET_INT": np.int32, "MET_ULONG_LONG": np.uint64, "MET_LONG_LONG": np.int64,
              "MET_FLOAT": np.float32, "MET_DOUBLE": np.float64}


@dataclass
class Volume3D:
    """Voxels as stored (z, y, x: x fastest, as ITK's buffer) plus ITK's physical frame:
    spacing and origin in (x, y, z) order, direction columns = index axes in LPS."""
    data: np.ndarray
    spacing: np.ndarray = field(default_factory=lambda: np.ones(3))
    origin: np.ndarray = field(default_factory=lambda: np.zeros(3))
    direction: np.ndarray = field(default_factory=lambda: np.eye(3))

    @property
    def size_xyz(self) -> tuple:
        return tuple(int(v) for v in self.data.shape[::-1])

    def idx_to_phys(self) -> np.ndarray:
        """ITKImagePhysicalPointTransformsAsEigen (lib/itk/xregITKBasicImageUtils.h:133-168):
        [D diag(spacing) | origin] as float32."""
        m = np.eye(4)
        m[:3, :3] = np.asarray(self.direction, np.float64) @ np.diag(np.asarray(self.spacing, np.float64))
        m[:3, 3] = self.origin
        return m.astype(np.float32)


def _open(path):
    path = os.fspath(path)
    return gzip.open(path, "rb") if path.endswith(".gz") else open(path, "rb")


def _quat_to_rot(b: float, c: float, d: float) -> np.ndarray:
    """nifti1.h quaternion convention: a = sqrt(1 - b^2 - c^2 - d^2) (clamped at 0)."""
    a2 = 1.0 - (b * b + c * c + d * d)
    a = np.sqrt(a2) if a2 > 1.0e-7 else 0.0
    if a == 0.0:  # 180-degree rotation: renormalise (b, c, d)
        n = np.sqrt(b * b + c * c + d * d)
        b, c, d = b / n, c / n, d / n
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a + c * c - b * b - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a + d * d - c * c - b * b]])


def read_nifti(path) -> Volume3D:
    """A 3-D NIfTI-1 volume (scalars; scl_slope applied when it is neither 0 nor 1 with a zero
    intercept, giving float32 data)."""
    with _open(path) as f:
        raw = f.read()
    if len(raw) < 348:
        raise ValueError("not a NIfTI-1 file (short header)")
    endian = "<" if struct.unpack("<i", raw[:4])[0] == 348 else ">"
    if struct.unpack(endian + "i", raw[:4])[0] != 348:
        raise ValueError("not a NIfTI-1 file (sizeof_hdr != 348)")
    magic = raw[344:348]
    if magic not in (b"n+1\x00", b"ni1\x00"):
        raise ValueError("not a NIfTI-1 file (magic)")
    if magic == b"ni1\x00":
        raise ValueError("two-file NIfTI (.hdr/.img) is not supported; use a single .nii")
    dim = struct.unpack(endian + "8h", raw[40:56])
    datatype, = struct.unpack(endian + "h", raw[70:72])
    pixdim = struct.unpack(endian + "8f", raw[76:108])
    vox_offset, scl_slope, scl_inter = struct.unpack(endian + "3f", raw[108:120])
    qform_code, sform_code = struct.unpack(endian + "2h", raw[252:256])
    qb, qc, qd, qx, qy, qz = struct.unpack(endian + "6f", raw[256:280])
    srow = np.array(struct.unpack(endian + "12f", raw[280:328]), np.float64).reshape(3, 4)
    nd = dim[0]
    if nd < 3 or any(dim[k] > 1 for k in range(4, nd + 1)):
        raise ValueError(f"only 3-D scalar volumes are supported (dim = {dim})")
    if datatype not in _NIFTI_DTYPES:
        raise ValueError(f"unsupported NIfTI datatype {datatype}")
    X, Y, Z = dim[1], dim[2], dim[3]
    dt = np.dtype(_NIFTI_DTYPES[datatype]).newbyteorder(endian)
    off = int(vox_offset)
    n = X * Y * Z
    data = np.frombuffer(raw, dtype=dt, count=n, offset=off).reshape(Z, Y, X).astype(dt.newbyteorder("="))
    if scl_slope != 0.0 and not (scl_slope == 1.0 and scl_inter == 0.0):
        data = (data.astype(np.float64) * scl_slope + scl_inter).astype(np.float32)
    sp = np.abs(np.array(pixdim[1:4], np.float64))
    if qform_code > 0:
        qfac = -1.0 if pixdim[0] < 0 else 1.0
        R = _quat_to_rot(qb, qc, qd)
        R[:, 2] *= qfac
        ras = R
        org = np.array([qx, qy, qz], np.float64)
    elif sform_code > 0:
        A = srow[:, :3]
        norms = np.linalg.norm(A, axis=0)
        ras = A / np.where(norms > 0, norms, 1.0)
        sp = np.where(norms > 0, norms, sp)
        org = srow[:, 3].copy()
    else:
        ras = np.eye(3)
        org = np.zeros(3)
    lps = np.diag([-1.0, -1.0, 1.0])
    return Volume3D(data, sp, lps @ org, lps @ ras)


def _rot_to_quat(R: np.ndarray):
    """Inverse of _quat_to_rot for a proper rotation (a >= 0)."""
    t = np.trace(R)
    if t > 0:
        s = 0.5 / np.sqrt(t + 1.0)
        a, b, c, d = 0.25 / s, (R[2, 1] - R[1, 2]) * s, (R[0, 2] - R[2, 0]) * s, (R[1, 0] - R[0, 1]) * s
    else:
        i = int(np.argmax(np.diag(R)))
        if i == 0:
            s = 2.0 * np.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2])
            a, b, c, d = (R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s
        elif i == 1:
            s = 2.0 * np.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2])
            a, b, c, d = (R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s
        else:
            s = 2.0 * np.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1])
            a, b, c, d = (R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s
    if a < 0:
        a, b, c, d = -a, -b, -c, -d
    return b, c, d


def write_nifti(path, vol: Volume3D) -> None:
    """A single-file NIfTI-1 with qform and sform both set to the volume's frame (RAS)."""
    data = np.ascontiguousarray(vol.data)
    key = data.dtype.newbyteorder("<").str[1:]
    if key not in _NIFTI_CODES:
        raise ValueError(f"unsupported voxel type {data.dtype}")
    lps = np.diag([-1.0, -1.0, 1.0])
    ras = lps @ np.asarray(vol.direction, np.float64)
    org = lps @ np.asarray(vol.origin, np.float64)
    qfac = 1.0
    R = ras.copy()
    if np.linalg.det(R) < 0:
        qfac = -1.0
        R[:, 2] *= -1.0
    qb, qc, qd = _rot_to_quat(R)
    sp = np.asarray(vol.spacing, np.float64)
    Z, Y, X = data.shape
    hdr = bytearray(352)
    struct.pack_into("<i", hdr, 0, 348)
    struct.pack_into("<8h", hdr, 40, 3, X, Y, Z, 1, 1, 1, 1)
    struct.pack_into("<2h", hdr, 70, _NIFTI_CODES[key], data.dtype.itemsize * 8)
    struct.pack_into("<8f", hdr, 76, qfac, sp[0], sp[1], sp[2], 1.0, 1.0, 1.0, 1.0)
    struct.pack_into("<3f", hdr, 108, 352.0, 1.0, 0.0)
    struct.pack_into("<2h", hdr, 252, 1, 1)
    struct.pack_into("<6f", hdr, 256, qb, qc, qd, org[0], org[1], org[2])
    A = ras @ np.diag(sp)
    for r in range(3):
        struct.pack_into("<4f", hdr, 280 + 16 * r, A[r, 0], A[r, 1], A[r, 2], org[r])
    hdr[344:348] = b"n+1\x00"
    payload = bytes(hdr) + data.astype(data.dtype.newbyteorder("<")).tobytes()
    p = os.fspath(path)
    with (gzip.open(p, "wb") if p.endswith(".gz") else open(p, "wb")) as f:
        f.write(payload)


def read_metaimage(path) -> Volume3D:
    """A 3-D MetaImage (.mha with LOCAL data, or .mhd with a data file beside it)."""
    path = os.fspath(path)
    with open(path, "rb") as f:
        blob = f.read()
    hdr, pos = {}, 0
    while True:
        nl = blob.find(b"\n", pos)
        if nl < 0:
            raise ValueError("MetaImage header has no ElementDataFile")
        line = blob[pos:nl].decode("latin-1").strip()
        pos = nl + 1
        if not line:
            continue
        k, _, v = line.partition("=")
        hdr[k.strip()] = v.strip()
        if k.strip() == "ElementDataFile":
            break
    if int(hdr.get("NDims", "3")) != 3:
        raise ValueError("only 3-D MetaImages are supported")
    if int(hdr.get("ElementNumberOfChannels", "1")) != 1:
        raise ValueError("only scalar MetaImages are supported")
    X, Y, Z = (int(v) for v in hdr["DimSize"].split())
    et = hdr["ElementType"]
    if et not in _MET_TYPES:
        raise ValueError(f"unsupported MetaImage element type {et}")
    msb = hdr.get("BinaryDataByteOrderMSB", hdr.get("ElementByteOrderMSB", "False")).lower() == "true"
    dt = np.dtype(_MET_TYPES[et]).newbyteorder(">" if msb else "<")
    src = hdr["ElementDataFile"]
    if src == "LOCAL":
        body = blob[pos:]
    else:
        with open(os.path.join(os.path.dirname(path), src), "rb") as f:
            body = f.read()
    if hdr.get("CompressedData", "False").lower() == "true":
        body = zlib.decompress(body)
    data = np.frombuffer(body, dtype=dt, count=X * Y * Z).reshape(Z, Y, X).astype(dt.newbyteorder("="))
    sp = np.array([float(v) for v in hdr.get("ElementSpacing", "1 1 1").split()])
    org = np.array([float(v) for v in hdr.get("Offset", hdr.get("Position", "0 0 0")).split()])
    tm = np.array([float(v) for v in hdr.get("TransformMatrix", hdr.get("Orientation", "1 0 0 0 1 0 0 0 1")).split()])
    direction = tm.reshape(3, 3).T  # row i of the listing is the direction of index axis i
    return Volume3D(data, sp, org, direction)


def write_metaimage(path, vol: Volume3D, compress: bool = False) -> None:
    """A .mha with LOCAL data."""
    data = np.ascontiguousarray(vol.data)
    rev = {np.dtype(v): k for k, v in _MET_TYPES.items()}
    if data.dtype.newbyteorder("=") not in rev:
        raise ValueError(f"unsupported voxel type {data.dtype}")
    Z, Y, X = data.shape
    body = data.astype(data.dtype.newbyteorder("<")).tobytes()
    if compress:
        body = zlib.compress(body)
    D = np.asarray(vol.direction, np.float64)
    lines = ["ObjectType = Image", "NDims = 3", "BinaryData = True", "BinaryDataByteOrderMSB = False",
             f"CompressedData = {'True' if compress else 'False'}",
             "TransformMatrix = " + " ".join(repr(float(v)) for v in D.T.reshape(-1)),
             "Offset = " + " ".join(repr(float(v)) for v in vol.origin),
             "ElementSpacing = " + " ".join(repr(float(v)) for v in vol.spacing),
             f"DimSize = {X} {Y} {Z}", f"ElementType = {rev[data.dtype.newbyteorder('=')]}",
             "ElementDataFile = LOCAL"]
    with open(os.fspath(path), "wb") as f:
        f.write(("\n".join(lines) + "\n").encode("latin-1"))
        f.write(body)


def read_volume(path) -> Volume3D:
    """By extension: .nii / .nii.gz (NIfTI-1), .mha / .mhd (MetaImage)."""
    p = os.fspath(path).lower()
    if p.endswith(".nii") or p.endswith(".nii.gz"):
        return read_nifti(path)
    if p.endswith(".mha") or p.endswith(".mhd"):
        return read_metaimage(path)
    raise ValueError(f"unknown volume file type: {path}")


def load_ct(ctx, path, hu: dict | None = None, storage: int | None = None, layout: int | None = None):
    """ReadITKImageFromDisk + HUToLinAtt + RayCaster::set_volume for the HIP path: the file's
    voxels go to the device as float32 with the file's index -> physical transform; hu (a dict,
    possibly empty, of hu_lower / mu_water / mu_air) converts Hounsfield units there, None
    keeps the voxels as attenuation. Returns (xreg_amd._xrh.Volume, Volume3D)."""
    from . import _xrh as X

    v = read_volume(path)
    dev = X.Volume(ctx, v.data.astype(np.float32), v.idx_to_phys(),
                   storage=X.XRH_F32 if storage is None else storage, layout=layout, hu=hu)
    return dev, v
