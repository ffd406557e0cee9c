"""HDF5 files of xReg: projection data, camera models, pose parameterisations and the
registration debug record.

The on-disk layout follows the reference's writers so files move both ways:

* ProjData (lib/file_formats/xregH5ProjDataIO.cpp:36-102, 338-420): group attribute
  ``xreg-type = "proj-data"``, scalar ``num-projs`` (uint64), then per projection a group
  ``proj-%03d`` holding ``img`` (optional), ``cam``, ``landmarks`` (optional),
  ``rot-to-pat-up`` (int32, optional), ``det-spacings-from-orig-meta`` (uint8, optional) and
  ``orig-dcm-meta`` (optional, kept as an opaque group of datasets).
* 2-D image (lib/hdf5/xregHDF5Internal.h:208-300, 505-547): attribute ``xreg-type =
  "image-2D"``, float matrices ``dir-mat`` (2x2), ``origin`` (2x1), ``spacing`` (2x1), and the
  dataset ``pixels`` shaped (rows, cols); with compression it is chunked (the reference's
  chunk-shrinking rule, at most 1 MiB per chunk) and deflated at level 9.
* Camera model (lib/file_formats/xregH5CamModelIO.cpp:29-90): ``num-rows`` / ``num-cols``
  (uint64), ``row-spacing`` / ``col-spacing`` (float), ``cam-coord-frame-type`` (fixed-length
  string), ``intrinsic`` (3x3 float) and ``extrinsic`` (4x4 float), row-major.
* SE3OptVars (lib/file_formats/xregH5SE3OptVarsIO.cpp:31-264): ``parameterization`` string
  plus the factor orders / axis flags.
* SingleRegiDebugResults (lib/regi/interfaces_2d_3d/xregIntensity2D3DRegiDebug.cpp:59-330).

Scalar types follow the reference's LookupH5DataType (lib/hdf5/xregHDF5.h:43-127): size_type
-> native unsigned long, bool -> unsigned char, CoordScalar -> float. Matrices are written
uncompressed where the reference passes ``compress = false`` and chunked + deflate(9) where it
uses the default (vectors and the iteration matrix of the debug record).

The HDF5 C library is loaded with ctypes: ``XREG_AMD_HDF5_LIB`` names it explicitly, else the
system search path, else the HDF5 1.10 shipped under /opt/conda/lib. Nothing here touches the
GPU.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os
from dataclasses import dataclass, field

import numpy as np

from . import geometry as G
from . import optvars as OV

# ------------------------------------------------------------------------------------------
# libhdf5 binding (HDF5 1.10 C API: hid_t is int64, herr_t/htri_t are int, hsize_t uint64)

hid_t = C.c_int64
herr_t = C.c_int
htri_t = C.c_int
hsize_t = C.c_uint64

H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5F_ACC_RDONLY = 0
H5F_ACC_RDWR = 1
H5F_ACC_TRUNC = 2
H5T_VARIABLE = C.c_size_t(-1).value
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5T_SGN_NONE = 0
H5D_CHUNKED = 2
H5Z_FILTER_DEFLATE = 1

_NATIVE_NAMES = ["FLOAT", "DOUBLE", "SCHAR", "UCHAR", "SHORT", "USHORT", "INT", "UINT", "LONG", "ULONG",
                 "LLONG", "ULLONG"]


class H5Error(RuntimeError):
    pass


class _GInfo(C.Structure):
    _fields_ = [("storage_type", C.c_int), ("nlinks", hsize_t), ("max_corder", C.c_int64), ("mounted", C.c_bool)]


class _Lib:
    def __init__(self, path: str):
        h = C.CDLL(path)
        self.h = h
        sig = {
            "H5open": (herr_t, []),
            "H5Eset_auto2": (herr_t, [hid_t, C.c_void_p, C.c_void_p]),
            "H5Fcreate": (hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]),
            "H5Fopen": (hid_t, [C.c_char_p, C.c_uint, hid_t]),
            "H5Fclose": (herr_t, [hid_t]),
            "H5Fflush": (herr_t, [hid_t, C.c_int]),
            "H5Gcreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t]),
            "H5Gopen2": (hid_t, [hid_t, C.c_char_p, hid_t]),
            "H5Gclose": (herr_t, [hid_t]),
            "H5Gget_info": (herr_t, [hid_t, C.POINTER(_GInfo)]),
            "H5Lexists": (htri_t, [hid_t, C.c_char_p, hid_t]),
            "H5Ldelete": (herr_t, [hid_t, C.c_char_p, hid_t]),
            "H5Lget_name_by_idx": (C.c_ssize_t, [hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, C.c_char_p,
                                                 C.c_size_t, hid_t]),
            "H5Ocopy": (herr_t, [hid_t, C.c_char_p, hid_t, C.c_char_p, hid_t, hid_t]),
            "H5Screate": (hid_t, [C.c_int]),
            "H5Screate_simple": (hid_t, [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
            "H5Sget_simple_extent_ndims": (C.c_int, [hid_t]),
            "H5Sget_simple_extent_dims": (C.c_int, [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
            "H5Sclose": (herr_t, [hid_t]),
            "H5Pcreate": (hid_t, [hid_t]),
            "H5Pset_chunk": (herr_t, [hid_t, C.c_int, C.POINTER(hsize_t)]),
            "H5Pset_deflate": (herr_t, [hid_t, C.c_uint]),
            "H5Pget_layout": (C.c_int, [hid_t]),
            "H5Pget_chunk": (C.c_int, [hid_t, C.c_int, C.POINTER(hsize_t)]),
            "H5Pget_nfilters": (C.c_int, [hid_t]),
            "H5Pget_filter2": (C.c_int, [hid_t, C.c_uint, C.POINTER(C.c_uint), C.POINTER(C.c_size_t),
                                         C.POINTER(C.c_uint), C.c_size_t, C.c_char_p, C.POINTER(C.c_uint)]),
            "H5Pclose": (herr_t, [hid_t]),
            "H5Dcreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
            "H5Dopen2": (hid_t, [hid_t, C.c_char_p, hid_t]),
            "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
            "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
            "H5Dget_space": (hid_t, [hid_t]),
            "H5Dget_type": (hid_t, [hid_t]),
            "H5Dget_create_plist": (hid_t, [hid_t]),
            "H5Dvlen_reclaim": (herr_t, [hid_t, hid_t, hid_t, C.c_void_p]),
            "H5Dclose": (herr_t, [hid_t]),
            "H5Tcopy": (hid_t, [hid_t]),
            "H5Tset_size": (herr_t, [hid_t, C.c_size_t]),
            "H5Tget_size": (C.c_size_t, [hid_t]),
            "H5Tget_class": (C.c_int, [hid_t]),
            "H5Tget_sign": (C.c_int, [hid_t]),
            "H5Tis_variable_str": (htri_t, [hid_t]),
            "H5Tclose": (herr_t, [hid_t]),
            "H5Acreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t]),
            "H5Aopen": (hid_t, [hid_t, C.c_char_p, hid_t]),
            "H5Aexists": (htri_t, [hid_t, C.c_char_p]),
            "H5Awrite": (herr_t, [hid_t, hid_t, C.c_void_p]),
            "H5Aread": (herr_t, [hid_t, hid_t, C.c_void_p]),
            "H5Aget_type": (hid_t, [hid_t]),
            "H5Aget_space": (hid_t, [hid_t]),
            "H5Aclose": (herr_t, [hid_t]),
        }
        for name, s in sig.items():
            fn = getattr(h, name)
            fn.restype, fn.argtypes = s
        if h.H5open() < 0:
            raise H5Error("H5open failed")
        h.H5Eset_auto2(0, None, None)  # errors come back as return codes, not printed
        self.T = {n: hid_t.in_dll(h, f"H5T_NATIVE_{n}_g").value for n in _NATIVE_NAMES}
        self.C_S1 = hid_t.in_dll(h, "H5T_C_S1_g").value
        self.DCPL = hid_t.in_dll(h, "H5P_CLS_DATASET_CREATE_ID_g").value

    def __getattr__(self, name):
        fn = getattr(self.h, name)
        if fn.restype in (hid_t, herr_t, htri_t, C.c_int, C.c_ssize_t):
            def call(*a):
                r = fn(*a)
                if r < 0:
                    raise H5Error(f"{name} failed")
                return r
            return call
        return fn


_LIB: _Lib | None = None


def lib() -> _Lib:
    """The loaded HDF5 library (raises H5Error with the places searched when it is absent)."""
    global _LIB
    if _LIB is None:
        cands = [os.environ.get("XREG_AMD_HDF5_LIB"), ctypes.util.find_library("hdf5"),
                 "/opt/conda/lib/libhdf5.so"]
        tried = []
        for c in cands:
            if not c:
                continue
            try:
                _LIB = _Lib(c)
                break
            except OSError as e:
                tried.append(f"{c}: {e}")
        if _LIB is None:
            raise H5Error("HDF5 C library not found (set XREG_AMD_HDF5_LIB); tried " + "; ".join(tried))
    return _LIB


def available() -> bool:
    try:
        lib()
        return True
    except H5Error:
        return False


# numpy dtype <-> native HDF5 type
_NP_TO_NATIVE = {
    np.dtype(np.float32): "FLOAT", np.dtype(np.float64): "DOUBLE", np.dtype(np.int8): "SCHAR",
    np.dtype(np.uint8): "UCHAR", np.dtype(np.int16): "SHORT", np.dtype(np.uint16): "USHORT",
    np.dtype(np.int32): "INT", np.dtype(np.uint32): "UINT", np.dtype(np.int64): "LONG",
    np.dtype(np.uint64): "ULONG",
}


def _native(dt) -> int:
    dt = np.dtype(dt)
    if dt not in _NP_TO_NATIVE:
        raise H5Error(f"no HDF5 native type for {dt}")
    return lib().T[_NP_TO_NATIVE[dt]]


def _np_dtype_of(type_id: int) -> np.dtype | None:
    L = lib()
    cls = L.H5Tget_class(type_id)
    size = L.H5Tget_size(type_id)
    if cls == H5T_FLOAT:
        return np.dtype({4: np.float32, 8: np.float64}[size])
    if cls == H5T_INTEGER:
        signed = L.H5Tget_sign(type_id) != H5T_SGN_NONE
        return np.dtype(f"{'i' if signed else 'u'}{size}")
    return None


def _dims(ids) -> tuple:
    L = lib()
    n = L.H5Sget_simple_extent_ndims(ids)
    if n == 0:
        return ()
    d = (hsize_t * n)()
    L.H5Sget_simple_extent_dims(ids, d, None)
    return tuple(int(x) for x in d)


# ------------------------------------------------------------------------------------------
# Group handles


class Group:
    """An open HDF5 file or group (H5::Group in the reference's helpers)."""

    def __init__(self, hid: int, is_file: bool = False):
        self.id = hid
        self._file = is_file

    # -- lifetime
    def close(self):
        if self.id is not None:
            L = lib()
            if self._file:
                L.H5Fflush(self.id, 1)
                L.H5Fclose(self.id)
            else:
                L.H5Gclose(self.id)
            self.id = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- structure
    def create_group(self, name: str) -> "Group":
        return Group(lib().H5Gcreate2(self.id, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT))

    def open_group(self, name: str) -> "Group":
        return Group(lib().H5Gopen2(self.id, name.encode(), H5P_DEFAULT))

    def __contains__(self, name: str) -> bool:
        # ObjectInGroupH5 (xregHDF5.cpp): every path component must exist
        cur = ""
        for part in name.split("/"):
            cur = f"{cur}/{part}" if cur else part
            if lib().H5Lexists(self.id, cur.encode(), H5P_DEFAULT) <= 0:
                return False
        return True

    def unlink(self, name: str):
        lib().H5Ldelete(self.id, name.encode(), H5P_DEFAULT)

    def names(self) -> list[str]:
        """Member names in name order (getObjnameByIdx of the reference)."""
        L = lib()
        info = _GInfo()
        L.H5Gget_info(self.id, C.byref(info))
        out = []
        for i in range(info.nlinks):
            n = L.H5Lget_name_by_idx(self.id, b".", 0, 0, i, None, 0, H5P_DEFAULT)
            buf = C.create_string_buffer(n + 1)
            L.H5Lget_name_by_idx(self.id, b".", 0, 0, i, buf, n + 1, H5P_DEFAULT)
            out.append(buf.value.decode())
        return out

    def is_group(self, name: str) -> bool:
        L = lib()
        try:
            g = L.H5Gopen2(self.id, name.encode(), H5P_DEFAULT)
        except H5Error:
            return False
        L.H5Gclose(g)
        return True

    # -- attributes
    def set_str_attr(self, key: str, val: str):
        """SetStringAttr (xregHDF5.cpp:59-78): fixed-length C string, scalar dataspace."""
        L = lib()
        b = val.encode()
        t = L.H5Tcopy(L.C_S1)
        L.H5Tset_size(t, max(1, len(b)))
        s = L.H5Screate(H5S_SCALAR)
        a = L.H5Acreate2(self.id, key.encode(), t, s, H5P_DEFAULT, H5P_DEFAULT)
        L.H5Awrite(a, t, C.create_string_buffer(b, max(1, len(b))))
        L.H5Aclose(a)
        L.H5Sclose(s)
        L.H5Tclose(t)

    def get_str_attr(self, key: str) -> str:
        L = lib()
        a = L.H5Aopen(self.id, key.encode(), H5P_DEFAULT)
        t = L.H5Aget_type(a)
        try:
            return _read_string(t, lambda mem, buf: L.H5Aread(a, mem, buf))
        finally:
            L.H5Tclose(t)
            L.H5Aclose(a)

    def has_attr(self, key: str) -> bool:
        return lib().H5Aexists(self.id, key.encode()) > 0

    # -- datasets
    def write_array(self, name: str, arr, compress: bool = False, chunk=None) -> None:
        """A numeric dataset of arr's shape and type (scalars use a scalar dataspace)."""
        L = lib()
        a = np.asarray(arr, order="C")  # (ascontiguousarray would turn a 0-d scalar into 1-d)
        if a.dtype == np.bool_:
            a = a.astype(np.uint8)
        t = _native(a.dtype)
        if a.ndim == 0:
            s = L.H5Screate(H5S_SCALAR)
        else:
            dims = (hsize_t * a.ndim)(*a.shape)
            s = L.H5Screate_simple(a.ndim, dims, None)
        dcpl = H5P_DEFAULT
        if compress and a.ndim > 0 and a.size > 0:
            dcpl = L.H5Pcreate(L.DCPL)
            ch = (hsize_t * a.ndim)(*(chunk if chunk is not None else a.shape))
            L.H5Pset_chunk(dcpl, a.ndim, ch)
            L.H5Pset_deflate(dcpl, 9)
        d = L.H5Dcreate2(self.id, name.encode(), t, s, H5P_DEFAULT, dcpl, H5P_DEFAULT)
        if a.size:
            L.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data)
        L.H5Dclose(d)
        if dcpl != H5P_DEFAULT:
            L.H5Pclose(dcpl)
        L.H5Sclose(s)

    def write_string(self, name: str, val: str) -> None:
        """WriteStringH5 with compress = false (xregHDF5.cpp:98-124)."""
        L = lib()
        b = val.encode()
        t = L.H5Tcopy(L.C_S1)
        L.H5Tset_size(t, max(1, len(b)))
        s = L.H5Screate(H5S_SCALAR)
        d = L.H5Dcreate2(self.id, name.encode(), t, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
        L.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, C.create_string_buffer(b, max(1, len(b))))
        L.H5Dclose(d)
        L.H5Sclose(s)
        L.H5Tclose(t)

    def read(self, name: str, dtype=None):
        """A dataset as a numpy array (0-d for scalars) converted to dtype (default: the file's
        own type), or a str for string datasets."""
        L = lib()
        d = L.H5Dopen2(self.id, name.encode(), H5P_DEFAULT)
        ft = L.H5Dget_type(d)
        s = L.H5Dget_space(d)
        try:
            if L.H5Tget_class(ft) == H5T_STRING:
                return _read_string(ft, lambda mem, buf: L.H5Dread(d, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf),
                                    reclaim=lambda mem, buf: L.H5Dvlen_reclaim(mem, s, H5P_DEFAULT, buf))
            dt = np.dtype(dtype) if dtype is not None else _np_dtype_of(ft)
            if dt is None:
                raise H5Error(f"unsupported HDF5 datatype for dataset {name}")
            out = np.empty(_dims(s), dtype=dt)
            if out.size:
                L.H5Dread(d, _native(dt), H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data)
            return out
        finally:
            L.H5Sclose(s)
            L.H5Tclose(ft)
            L.H5Dclose(d)

    def scalar(self, name: str, dtype):
        """ReadSingleScalarH5*: the one value of a scalar (or one-element) dataset."""
        a = self.read(name, dtype)
        if a.size != 1:
            raise H5Error(f"dataset {name} holds {a.size} values, not one")
        return a.reshape(-1)[0]

    def dataset_dtype(self, name: str) -> np.dtype | None:
        L = lib()
        d = L.H5Dopen2(self.id, name.encode(), H5P_DEFAULT)
        ft = L.H5Dget_type(d)
        try:
            return _np_dtype_of(ft)
        finally:
            L.H5Tclose(ft)
            L.H5Dclose(d)

    def storage(self, name: str) -> dict:
        """Chunk shape and deflate level of a dataset (None when contiguous / uncompressed)."""
        L = lib()
        d = L.H5Dopen2(self.id, name.encode(), H5P_DEFAULT)
        p = L.H5Dget_create_plist(d)
        s = L.H5Dget_space(d)
        try:
            out = {"chunk": None, "deflate": None}
            if L.H5Pget_layout(p) == H5D_CHUNKED:
                n = L.H5Sget_simple_extent_ndims(s)
                ch = (hsize_t * n)()
                L.H5Pget_chunk(p, n, ch)
                out["chunk"] = tuple(int(x) for x in ch)
            for i in range(L.H5Pget_nfilters(p)):
                flags, nelem, namebuf = C.c_uint(), C.c_size_t(1), C.create_string_buffer(64)
                cd = (C.c_uint * 1)()
                fc = C.c_uint()
                fid = L.H5Pget_filter2(p, i, C.byref(flags), C.byref(nelem), cd, 64, namebuf, C.byref(fc))
                if fid == H5Z_FILTER_DEFLATE:
                    out["deflate"] = int(cd[0])
            return out
        finally:
            L.H5Sclose(s)
            L.H5Pclose(p)
            L.H5Dclose(d)

    def copy(self, name: str, dst: "Group", dst_name: str | None = None):
        lib().H5Ocopy(self.id, name.encode(), dst.id, (dst_name or name).encode(), H5P_DEFAULT, H5P_DEFAULT)


def _read_string(ftype: int, read, reclaim=None) -> str:
    L = lib()
    if L.H5Tis_variable_str(ftype) > 0:
        mem = L.H5Tcopy(L.C_S1)
        L.H5Tset_size(mem, H5T_VARIABLE)
        p = C.c_char_p()
        read(mem, C.byref(p))
        val = (p.value or b"").decode()
        if reclaim is not None:
            reclaim(mem, C.byref(p))
        L.H5Tclose(mem)
        return val
    n = L.H5Tget_size(ftype)
    mem = L.H5Tcopy(L.C_S1)
    L.H5Tset_size(mem, n + 1)
    buf = C.create_string_buffer(n + 1)
    read(mem, buf)
    L.H5Tclose(mem)
    return buf.value.decode()


def create_file(path: str) -> Group:
    """H5File(path, H5F_ACC_TRUNC)."""
    return Group(lib().H5Fcreate(os.fspath(path).encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), is_file=True)


def open_file(path: str, writable: bool = False) -> Group:
    return Group(lib().H5Fopen(os.fspath(path).encode(), H5F_ACC_RDWR if writable else H5F_ACC_RDONLY,
                               H5P_DEFAULT), is_file=True)


# ------------------------------------------------------------------------------------------
# Helpers with the reference's types


def write_ulong(g: Group, name: str, v: int):
    """WriteSingleScalarH5 of a size_type (native unsigned long)."""
    g.write_array(name, np.uint64(v))


def write_bool(g: Group, name: str, v: bool):
    """WriteSingleScalarH5 of a bool: stored as unsigned char (xregHDF5.cpp:196-201)."""
    g.write_array(name, np.uint8(bool(v)))


def write_matrix(g: Group, name: str, m, compress: bool = True, dtype=np.float32):
    """WriteMatrixH5 (xregHDF5Internal.h:106-176): a 2-D row-major dataset; with compression,
    chunks of at most 1024 / sizeof(scalar) elements per dimension, deflate 9."""
    a = np.asarray(m, dtype=dtype)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    per = 1024 // a.dtype.itemsize
    chunk = (min(a.shape[0], per), min(a.shape[1], per))
    g.write_array(name, a, compress=compress, chunk=chunk)


def write_vector(g: Group, name: str, v, dtype, compress: bool = True):
    """WriteVectorH5 (xregHDF5Internal.h:49-88): 1-D, chunk <= 1 MiB, deflate 9."""
    a = np.asarray(v, dtype=dtype).reshape(-1)
    chunk = (min(a.size, (1 << 20) // a.dtype.itemsize),)
    g.write_array(name, a, compress=compress, chunk=chunk)


def image_chunk_dims(shape, itemsize: int) -> tuple:
    """Chunk dimensions of WriteNDImageH5Helper (xregHDF5Internal.h:258-291): starting from the
    whole image, the first dimension that is not 1 is decremented until the chunk holds at most
    1 MiB. Each dimension's decrements end at the largest size that fits (or 1), computed here
    directly."""
    ch = [int(x) for x in shape]
    cap = (1 << 20) // itemsize
    for i in range(len(ch)):
        if int(np.prod(ch)) <= cap:
            break
        rest = int(np.prod(ch[i + 1:]))
        ch[i] = max(1, min(ch[i], cap // rest))
    return tuple(ch)


@dataclass
class Image2D:
    """itk::Image<T,2> as the files hold it: pixels (rows, cols) plus ITK metadata in (x, y)
    = (col, row) order."""
    pixels: np.ndarray
    spacing: np.ndarray = field(default_factory=lambda: np.ones(2, np.float32))
    origin: np.ndarray = field(default_factory=lambda: np.zeros(2, np.float32))
    direction: np.ndarray = field(default_factory=lambda: np.eye(2, dtype=np.float32))


def write_image(g: Group, img: Image2D, compress: bool = True):
    """WriteImageH5 for a 2-D image (xregHDF5Internal.h:208-300)."""
    g.set_str_attr("xreg-type", f"image-{img.pixels.ndim}D")
    write_matrix(g, "dir-mat", img.direction, compress=False)
    write_matrix(g, "origin", np.asarray(img.origin).reshape(-1, 1), compress=False)
    write_matrix(g, "spacing", np.asarray(img.spacing).reshape(-1, 1), compress=False)
    px = np.ascontiguousarray(img.pixels)
    g.write_array("pixels", px, compress=compress, chunk=image_chunk_dims(px.shape, px.dtype.itemsize))


def read_image(g: Group, dtype=None) -> Image2D:
    """ReadNDImageH5Helper (xregHDF5Internal.h:505-547); pixels converted to dtype if given."""
    return Image2D(pixels=g.read("pixels", dtype),
                   spacing=g.read("spacing", np.float32).reshape(-1),
                   origin=g.read("origin", np.float32).reshape(-1),
                   direction=g.read("dir-mat", np.float32))


_FRAME_STR = {G.FRAME_POS_Z: "origin-at-focal-pt-det-pos-z", G.FRAME_NEG_Z: "origin-at-focal-pt-det-neg-z",
              G.FRAME_ON_DETECTOR: "origin-on-det"}


def write_cam_model(g: Group, cam) -> None:
    """WriteCamModelH5 (xregH5CamModelIO.cpp:29-58). cam is a Camera carrying K and E."""
    g.set_str_attr("xreg-type", "cam-model")
    write_ulong(g, "num-rows", cam.num_det_rows)
    write_ulong(g, "num-cols", cam.num_det_cols)
    g.write_array("row-spacing", np.float32(cam.det_row_spacing))
    g.write_array("col-spacing", np.float32(cam.det_col_spacing))
    g.write_string("cam-coord-frame-type", _FRAME_STR.get(int(cam.coord_frame), "origin-on-det"))
    write_matrix(g, "intrinsic", cam.K, compress=False)
    write_matrix(g, "extrinsic", cam.E, compress=False)


def read_cam_model(g: Group):
    """ReadCamModelH5 (xregH5CamModelIO.cpp:70-95): frame from the string (anything unknown is
    origin-on-detector), then CameraModel::setup(intrinsic, extrinsic, rows, cols, rs, cs)."""
    s = g.read("cam-coord-frame-type")
    frame = {v: k for k, v in _FRAME_STR.items()}.get(s, G.FRAME_ON_DETECTOR)
    return G.camera_ke(g.read("intrinsic", np.float32), g.read("extrinsic", np.float32),
                       int(g.scalar("num-rows", np.uint64)), int(g.scalar("num-cols", np.uint64)),
                       float(g.scalar("row-spacing", np.float32)), float(g.scalar("col-spacing", np.float32)),
                       coord_frame=frame)


# ------------------------------------------------------------------------------------------
# ProjData


@dataclass
class ProjData:
    """xreg::ProjData (lib/image/xregProjData.h:52-82). img is None when the pixels were not
    read (or the projection has none); landmarks map names to 2-D index points."""
    cam: object
    img: Image2D | None = None
    landmarks: dict = field(default_factory=dict)
    rot_to_pat_up: int | None = None
    det_spacings_from_orig_meta: bool | None = None
    orig_dcm_meta: dict | None = None


ROT_TO_PAT_UP = (0, 90, 180, 270)  # ProjDataRotToPatUp (xregProjData.h:40-46)
SCALAR_TYPES = {np.dtype(np.float32): "float32", np.dtype(np.uint16): "uint16", np.dtype(np.uint8): "uint8"}


def _write_dict(g: Group, d: dict):
    for k, v in d.items():
        if isinstance(v, str):
            g.write_string(k, v)
        else:
            g.write_array(k, np.asarray(v))


def _read_dict(g: Group) -> dict:
    out = {}
    for n in g.names():
        if not g.is_group(n):
            out[n] = g.read(n)
    return out


def write_proj_data_group(g: Group, projs: list[ProjData], compress: bool = True) -> None:
    """WriteProjDataH5 (xregH5ProjDataIO.cpp:36-102)."""
    g.set_str_attr("xreg-type", "proj-data")
    write_ulong(g, "num-projs", len(projs))
    for i, pd in enumerate(projs):
        pg = g.create_group(f"proj-{i:03d}")
        if pd.img is not None:
            img = pd.img if isinstance(pd.img, Image2D) else Image2D(
                np.asarray(pd.img), spacing=np.array([pd.cam.det_col_spacing, pd.cam.det_row_spacing], np.float32))
            if img.pixels.dtype not in SCALAR_TYPES:
                raise H5Error(f"proj data pixels must be float32, uint16 or uint8, not {img.pixels.dtype}")
            with pg.create_group("img") as ig:
                write_image(ig, img, compress)
        with pg.create_group("cam") as cg:
            write_cam_model(cg, pd.cam)
        if pd.landmarks:
            with pg.create_group("landmarks") as lg:
                for name, pt in pd.landmarks.items():
                    write_matrix(lg, name, np.asarray(pt, np.float32).reshape(-1, 1), compress=False)
        if pd.rot_to_pat_up is not None:
            if int(pd.rot_to_pat_up) not in ROT_TO_PAT_UP:
                raise H5Error("rot_to_pat_up must be 0, 90, 180 or 270")
            pg.write_array("rot-to-pat-up", np.int32(pd.rot_to_pat_up))
        if pd.det_spacings_from_orig_meta is not None:
            write_bool(pg, "det-spacings-from-orig-meta", pd.det_spacings_from_orig_meta)
        if pd.orig_dcm_meta is not None:
            with pg.create_group("orig-dcm-meta") as mg:
                mg.set_str_attr("meta-type", "dicom")
                _write_dict(mg, pd.orig_dcm_meta)
        pg.close()


def read_proj_data_group(g: Group, read_pixels: bool = True, dtype=None) -> list[ProjData]:
    """ReadProjDataH5* (xregH5ProjDataIO.cpp:338-420). dtype converts the pixels the way the
    typed readers do (HDF5 converts on read); None keeps the file's own type."""
    n = int(g.scalar("num-projs", np.uint64))
    out = []
    for i in range(n):
        with g.open_group(f"proj-{i:03d}") as pg:
            with pg.open_group("cam") as cg:
                pd = ProjData(cam=read_cam_model(cg))
            if read_pixels and "img" in pg:
                with pg.open_group("img") as ig:
                    pd.img = read_image(ig, dtype)
            if "landmarks" in pg:
                with pg.open_group("landmarks") as lg:
                    pd.landmarks = {k: lg.read(k, np.float32).reshape(-1) for k in lg.names()}
            if "rot-to-pat-up" in pg:
                pd.rot_to_pat_up = int(pg.scalar("rot-to-pat-up", np.int32))
            if "det-spacings-from-orig-meta" in pg:
                pd.det_spacings_from_orig_meta = bool(pg.scalar("det-spacings-from-orig-meta", np.uint8))
            if "orig-dcm-meta" in pg:
                with pg.open_group("orig-dcm-meta") as mg:
                    if mg.get_str_attr("meta-type") != "dicom":
                        raise H5Error("orig-dcm-meta is not DICOM metadata")
                    pd.orig_dcm_meta = _read_dict(mg)
            out.append(pd)
    return out


def write_proj_data(path: str, projs: list[ProjData], compress: bool = True) -> None:
    """WriteProjDataH5ToDisk (xregH5ProjDataIO.cpp:221-249)."""
    with create_file(path) as f:
        write_proj_data_group(f, projs, compress)


def read_proj_data(path: str, read_pixels: bool = True, dtype=None) -> list[ProjData]:
    """ReadProjDataH5{F32,U16,U8}FromDisk (dtype float32 / uint16 / uint8, or None)."""
    with open_file(path) as f:
        return read_proj_data_group(f, read_pixels, dtype)


def read_single_img(path: str, proj_idx: int, dtype=None) -> Image2D:
    """ReadSingleImgFromProjDataH5*FromDisk (xregH5ProjDataIO.cpp:404-412)."""
    with open_file(path) as f:
        n = int(f.scalar("num-projs", np.uint64))
        if not 0 <= proj_idx < n:
            raise H5Error(f"projection index {proj_idx} out of range (num-projs {n})")
        with f.open_group(f"proj-{proj_idx:03d}/img") as ig:
            return read_image(ig, dtype)


def read_cam_models(path: str) -> list:
    """ReadCamModelsFromProjDataFromDisk."""
    return [pd.cam for pd in read_proj_data(path, read_pixels=False)]


def proj_data_scalar_type(path: str) -> str:
    """GetProjDataScalarTypeH5 (xregH5ProjDataIO.cpp:487-526): the pixel type of every
    projection, which must agree ("float32", "uint16" or "uint8")."""
    with open_file(path) as f:
        n = int(f.scalar("num-projs", np.uint64))
        if n == 0:
            raise H5Error("no projections")
        kinds = set()
        for i in range(n):
            with f.open_group(f"proj-{i:03d}/img") as ig:
                dt = ig.dataset_dtype("pixels")
            if dt not in SCALAR_TYPES:
                raise H5Error("unsupported HDF5 datatype for proj data!")
            kinds.add(SCALAR_TYPES[dt])
        if len(kinds) != 1:
            raise H5Error("Inconsistent proj data pixel types!")
        return kinds.pop()


def copy_proj_data(src_path: str, dst_path: str) -> None:
    """ReadProjDataFromH5AndWriteToDisk / CopyProjDataH5 (xregH5ProjDataIO.cpp:104-141)."""
    with open_file(src_path) as s, create_file(dst_path) as d:
        d.set_str_attr("xreg-type", "proj-data")
        s.copy("num-projs", d)
        for i in range(int(s.scalar("num-projs", np.uint64))):
            s.copy(f"proj-{i:03d}", d)


def add_lands_to_proj_data(path: str, lands: dict, proj_idx: int, delete_existing: bool = False) -> None:
    """AddLandsToProjDataH5 (xregH5ProjDataIO.cpp:36-53, 660-689) on a file on disk."""
    with open_file(path, writable=True) as f:
        with f.open_group(f"proj-{proj_idx:03d}") as pg:
            if delete_existing and "landmarks" in pg:
                pg.unlink("landmarks")
            if lands:
                lg = pg.open_group("landmarks") if "landmarks" in pg else pg.create_group("landmarks")
                with lg:
                    for name, pt in lands.items():
                        if name in lg:
                            lg.unlink(name)
                        write_matrix(lg, name, np.asarray(pt, np.float32).reshape(-1, 1), compress=False)


class DeferredProjReader:
    """DeferredProjReader (xregH5ProjDataIO.cpp:560-655): cameras up front, pixels on demand,
    optionally cached."""

    def __init__(self, path: str, cache_imgs: bool = False):
        self.path = os.fspath(path)
        self.cache_imgs = cache_imgs
        self.scalar_type_on_disk = proj_data_scalar_type(self.path)
        self.proj_data = read_proj_data(self.path, read_pixels=False)

    def num_projs_on_disk(self) -> int:
        return len(self.proj_data)

    def read_proj(self, idx: int, dtype=np.float32) -> Image2D:
        pd = self.proj_data[idx]
        if self.cache_imgs and pd.img is not None and pd.img.pixels.dtype == np.dtype(dtype):
            return pd.img
        img = read_single_img(self.path, idx, dtype)
        if self.cache_imgs:
            pd.img = img
        return img


# ------------------------------------------------------------------------------------------
# SE3OptVars and the registration debug record


def write_se3_opt_vars(g: Group, ov) -> None:
    """WriteSE3OptVarsH5 (xregH5SE3OptVarsIO.cpp:99-264)."""
    write_ulong(g, "num-params", ov.num_params)
    if isinstance(ov, OV.SE3OptVarsEuler):
        g.write_string("parameterization", "se3-euler")
        pos = {p: k for k, (_, p) in enumerate(ov.order)}
        for p, n in enumerate(["rot-x-order", "rot-y-order", "rot-z-order", "trans-x-order", "trans-y-order",
                               "trans-z-order"]):
            write_ulong(g, n, pos[p])
    elif isinstance(ov, OV.SE3OptVarsLieAlg):
        g.write_string("parameterization", "se3-lie-alg")
    elif isinstance(ov, OV.SE3OptVarsTransOnly):
        if len(ov.axes) == 1:
            g.write_string("parameterization", f"trans-{'xyz'[ov.axes[0]]}-only")
        else:
            g.write_string("parameterization", "translation-only")
            for a in range(3):
                write_bool(g, f"trans-{'xyz'[a]}", a in ov.axes)
    elif isinstance(ov, OV.SO3OptVarsLieAlg):
        g.write_string("parameterization", "so3-lie-alg")
    elif isinstance(ov, OV.SO3OptVarsEuler):
        g.write_string("parameterization", "rot-only-euler")
        pos = {p: k for k, (_, p) in enumerate(ov.order)}
        for p, n in enumerate(["rot-x-order", "rot-y-order", "rot-z-order"]):
            write_ulong(g, n, pos[p])
    elif isinstance(ov, OV.SO3OptVarsOnly):
        g.write_string("parameterization", f"rot-{'xyz'[ov.axis]}-only")
    else:
        raise H5Error(f"unsupported SE3 parameterization {type(ov).__name__}")


def read_se3_opt_vars(g: Group):
    """ReadSE3OptVarsH5 (xregH5SE3OptVarsIO.cpp:31-97)."""
    s = g.read("parameterization")
    u = lambda n: int(g.scalar(n, np.uint64))  # noqa: E731
    b = lambda n: bool(g.scalar(n, np.uint8))  # noqa: E731
    if s == "se3-euler":
        return OV.SE3OptVarsEuler(u("rot-x-order"), u("rot-y-order"), u("rot-z-order"), u("trans-x-order"),
                                  u("trans-y-order"), u("trans-z-order"))
    if s == "se3-lie-alg":
        return OV.SE3OptVarsLieAlg()
    if s == "translation-only":
        return OV.SE3OptVarsTransOnly(b("trans-x"), b("trans-y"), b("trans-z"))
    if s in ("trans-x-only", "trans-y-only", "trans-z-only"):
        a = "xyz".index(s[6])
        return OV.SE3OptVarsTransOnly(a == 0, a == 1, a == 2)
    if s == "so3-lie-alg":
        return OV.SO3OptVarsLieAlg()
    if s == "rot-only-euler":
        return OV.SO3OptVarsEuler(u("rot-x-order"), u("rot-y-order"), u("rot-z-order"))
    if s in ("rot-x-only", "rot-y-only", "rot-z-only"):
        return OV.SO3OptVarsOnly("xyz".index(s[4]))
    raise H5Error(f"unsupported SE3 parameterization: {s}")


@dataclass
class RegiDebugResults:
    """SingleRegiDebugResults (xregIntensity2D3DRegiDebug.h:45-141). iter_vars[obj] is
    (iterations, num_params). opt_aux: the optimiser's record (regi.run_cmaes /
    regi.run_exhaustive put it in RegiResult.extra["opt_aux"]) or None; sims_aux: one metric
    record per view (regi.Objective.sims_aux) or None per view."""
    se3_params: object
    vols_used: list
    init_poses: list
    final_poses: list
    inter_frames: list
    inter_frames_wrt_vol: list
    iter_vars: list
    sims: np.ndarray
    projs_used: list = field(default_factory=list)
    static_vols: list = field(default_factory=list)
    static_vol_poses: list = field(default_factory=list)
    do_not_use_static_vol_heuristic: bool = False
    regi_time_secs: float | None = None
    opt_aux: dict | None = None
    sims_aux: list = field(default_factory=list)


def write_opt_aux(g: Group, aux: dict) -> None:
    """The optimiser's "opt-aux" group: Intensity2D3DRegiCMAES::OptAux::write
    (xregIntensity2D3DRegiCMAES.cpp:437-509) or Intensity2D3DRegiExhaustive::OptAux::write
    (xregIntensity2D3DRegiExhaustive.cpp:397-440), same names and on-disk types."""
    with g.create_group("opt-aux") as ag:
        ag.write_string("name", aux["name"])
        if aux["name"] == "CMA-ES":
            nits = len(aux["cov_mats"])
            write_ulong(ag, "num-iterations", nits)
            write_ulong(ag, "pop-size", aux["pop_size"])
            dims = list(aux.get("obj_dims") or [np.asarray(aux["pop_params"][0]).shape[1] if nits else 0])
            with ag.create_group("all-pop-params") as pg:
                for i, P in enumerate(aux["pop_params"]):
                    with pg.create_group(f"iter-{i:03d}") as ig:
                        o = 0
                        for oi, d in enumerate(dims):
                            write_matrix(ig, f"obj-{oi:03d}-pop", np.asarray(P)[:, o:o + d])
                            o += d
            if nits:
                write_matrix(ag, "all-sim-vals", np.stack([np.asarray(v) for v in aux["sim_vals"]]))
            with ag.create_group("covariances") as cg:
                for i, Cm in enumerate(aux["cov_mats"]):
                    write_matrix(cg, f"iter-{i:03d}", Cm)
            if len(aux.get("num_rejects", [])):
                write_vector(ag, "all-num-rejects", aux["num_rejects"], np.uint64)
        elif aux["name"] == "Exhaustive":
            vols = aux.get("cam_wrt_vols") or []
            if vols and len(vols[0]):
                with ag.create_group("cam-wrt-vols") as cwg:
                    for vi, poses in enumerate(vols):
                        with cwg.create_group(f"vol-{vi:02d}") as vg:
                            for i, T in enumerate(np.asarray(poses).reshape(-1, 4, 4)):
                                write_matrix(vg, f"{i:06d}", T)
            if aux.get("has_min"):
                ag.write_array("min-sim-val", np.float32(aux["min_sim_val"]))
                if aux.get("has_all_sim_vals"):
                    write_ulong(ag, "min-sim-val-idx", aux["min_sim_val_idx"])
            if aux.get("has_all_sim_vals"):
                write_vector(ag, "all-sim-vals", aux["all_sim_vals"], np.float32)
        else:
            raise H5Error(f"unknown optimiser record {aux['name']!r}")


def _write_patch_ncc_aux(g: Group, rec: dict) -> None:
    """ImgSimMetric2DPatchNCCCPU::SimAux::write (xregImgSimMetric2DPatchNCCCPU.cpp:443-485): one
    group per compute call with the patch bounds, weights and the indices of the patches used."""
    g.write_string("sim-aux-type", "patch-ncc-aux")
    p = rec["patches"]
    for i in range(int(rec["calls"])):
        with g.create_group(f"{i:04d}") as cg:
            write_vector(cg, "start-rows", p["start_rows"], np.uint64)
            write_vector(cg, "start-cols", p["start_cols"], np.uint64)
            write_vector(cg, "stop-rows", p["stop_rows"], np.uint64)
            write_vector(cg, "stop-cols", p["stop_cols"], np.uint64)
            write_vector(cg, "weights", p["weights"], np.float32)
            write_vector(cg, "patch-inds", rec["patch_inds"], np.uint64)


def write_sim_aux(g: Group, rec: dict) -> None:
    """One metric's record: patch NCC, or patch gradient-NCC with its x and y records
    (ImgSimMetric2DPatchGradNCCCPU::SimAux::write, xregImgSimMetric2DPatchGradNCCCPU.cpp:411-420)."""
    if rec["type"] == "patch-grad-ncc-aux":
        g.write_string("sim-aux-type", "patch-grad-ncc-aux")
        with g.create_group("x") as xg:
            _write_patch_ncc_aux(xg, rec["x"])
        with g.create_group("y") as yg:
            _write_patch_ncc_aux(yg, rec["y"])
    elif rec["type"] == "patch-ncc-aux":
        _write_patch_ncc_aux(g, rec)
    else:
        raise H5Error(f"unknown metric record {rec['type']!r}")


def write_regi_debug_group(g: Group, r: RegiDebugResults) -> None:
    """WriteSingleRegiDebugResultsH5 (xregIntensity2D3DRegiDebug.cpp:59-200)."""
    with g.create_group("se3-parameterization") as sg:
        write_se3_opt_vars(sg, r.se3_params)
    if len(r.projs_used):
        write_vector(g, "projs-used", r.projs_used, np.uint64)
    with g.create_group("static-vols") as svg:
        write_bool(svg, "no-static-vol-heuristic", r.do_not_use_static_vol_heuristic)
        if len(r.static_vols) != len(r.static_vol_poses):
            raise H5Error("static_vols and static_vol_poses differ in length")
        if r.static_vols:
            write_vector(svg, "vols", r.static_vols, np.uint64)
            with svg.create_group("poses") as pg:
                for k, T in enumerate(r.static_vol_poses):
                    write_matrix(pg, f"{k:03d}", T, compress=True)
    nv = len(r.init_poses)
    if not (nv == len(r.final_poses) == len(r.iter_vars) == len(r.vols_used)):
        raise H5Error("per-object lists differ in length")
    if nv == 0:
        return
    write_ulong(g, "num-objs", nv)
    nits = int(np.asarray(r.iter_vars[0]).shape[0])
    if nits != len(r.sims):
        raise H5Error("iteration count differs from the number of similarity values")
    write_ulong(g, "num-iterations", nits)
    gm, ip, fp, itf, pv = (g.create_group(n) for n in
                           ["global-obj-index-map", "init-poses", "final-poses", "inter-frames", "iter-pose-vars"])
    for k in range(nv):
        name = f"obj-{k:03d}"
        write_matrix(ip, name, r.init_poses[k], compress=False)
        write_matrix(fp, name, r.final_poses[k], compress=False)
        write_matrix(itf, name, r.inter_frames[k], compress=False)
        write_bool(itf, f"{name}-wrt-vol", r.inter_frames_wrt_vol[k])
        write_ulong(gm, name, r.vols_used[k])
        if nits:
            write_matrix(pv, name, np.asarray(r.iter_vars[k], np.float32).reshape(nits, -1), compress=True)
    for h in (gm, ip, fp, itf, pv):
        h.close()
    if r.regi_time_secs is not None:
        g.write_array("regi-elapsed-seconds", np.float64(r.regi_time_secs))
    if nits:
        write_vector(g, "sim-vals", r.sims, np.float32)
    if r.opt_aux is not None:
        write_opt_aux(g, r.opt_aux)
    if any(a is not None for a in r.sims_aux):
        with g.create_group("sims-aux") as sag:
            for si, rec in enumerate(r.sims_aux):
                if rec is not None:
                    with sag.create_group(f"{si:03d}") as sg:
                        write_sim_aux(sg, rec)


def read_regi_debug_group(g: Group) -> RegiDebugResults:
    """ReadSingleRegiDebugResultsH5 (xregIntensity2D3DRegiDebug.cpp:206-330), same fields."""
    with g.open_group("se3-parameterization") as sg:
        se3 = read_se3_opt_vars(sg)
    r = RegiDebugResults(se3_params=se3, vols_used=[], init_poses=[], final_poses=[], inter_frames=[],
                         inter_frames_wrt_vol=[], iter_vars=[], sims=np.zeros(0, np.float32))
    if "projs-used" in g:
        r.projs_used = [int(x) for x in g.read("projs-used", np.uint64)]
    if "static-vols" in g:
        with g.open_group("static-vols") as svg:
            r.do_not_use_static_vol_heuristic = bool(svg.scalar("no-static-vol-heuristic", np.uint8))
            if "vols" in svg:
                r.static_vols = [int(x) for x in svg.read("vols", np.uint64)]
                with svg.open_group("poses") as pg:
                    r.static_vol_poses = [pg.read(f"{k:03d}", np.float32) for k in range(len(r.static_vols))]
    if "num-objs" not in g:
        return r
    nv = int(g.scalar("num-objs", np.uint64))
    nits = int(g.scalar("num-iterations", np.uint64))
    with g.open_group("global-obj-index-map") as gm, g.open_group("init-poses") as ip, \
            g.open_group("final-poses") as fp, g.open_group("inter-frames") as itf, \
            g.open_group("iter-pose-vars") as pv:
        for k in range(nv):
            name = f"obj-{k:03d}"
            r.vols_used.append(int(gm.scalar(name, np.uint64)))
            r.init_poses.append(ip.read(name, np.float32))
            r.final_poses.append(fp.read(name, np.float32))
            r.inter_frames.append(itf.read(name, np.float32))
            r.inter_frames_wrt_vol.append(bool(itf.scalar(f"{name}-wrt-vol", np.uint8)))
            r.iter_vars.append(pv.read(name, np.float32) if nits else np.zeros((0, se3.num_params), np.float32))
    if "regi-elapsed-seconds" in g:
        r.regi_time_secs = float(g.scalar("regi-elapsed-seconds", np.float64))
    if nits:
        r.sims = g.read("sim-vals", np.float32)
    return r


def write_regi_debug(path: str, r: RegiDebugResults) -> None:
    with create_file(path) as f:
        write_regi_debug_group(f, r)


def read_regi_debug(path: str) -> RegiDebugResults:
    with open_file(path) as f:
        return read_regi_debug_group(f)
