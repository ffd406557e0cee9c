"""The C-ABI library loads and exports every entry point include/xrh.h declares (CPU only: no
compute call is made)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "xrh.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xrh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ["xrh_ctx_create", "xrh_volume_create", "xrh_line_int", "xrh_fixed_create",
                     "xrh_sim_compute", "xrh_batch_create", "xrh_batch_eval", "xrh_last_error"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    from xreg_amd import _xrh as X

    lib = ctypes.CDLL(str(X.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes binding covers the full surface too
    assert set(declared_functions()) == set(X.SIGNATURES)


def test_version_and_error_without_gpu():
    from xreg_amd import _xrh as X

    lib = X.lib()
    assert lib.xrh_abi_version() == 1
    assert isinstance(lib.xrh_last_error(), bytes)


def test_struct_layouts_match_header():
    from xreg_amd import _xrh as X

    # xrh_camera: 3 int32 + float + 9 + 16 + 3 + 2 floats
    assert ctypes.sizeof(X.Camera) == 4 * (4 + 9 + 16 + 3 + 2)
    assert ctypes.sizeof(X.SimParams) == 20
    assert X.View.cam.offset == 16


def test_volume_slice_past_24bit_index_reach_is_refused_without_gpu():
    """The DRR kernels multiply indices by the padded pitches in signed 24-bit arithmetic, so a
    volume whose padded slice (X+2)*(Y+2) reaches 2^23 voxels is refused (XRH_ERR_UNSUPPORTED)
    before any device work; the argument check runs ahead of the context check, so no GPU is
    needed to see it."""
    import numpy as np

    from xreg_amd import _xrh as X

    lib = X.lib()
    vox = np.zeros(4, np.float32)
    i2p = np.eye(4, dtype=np.float32).reshape(16)
    out = ctypes.c_void_p()
    for size, want in (((2900, 2900, 20), X.XRH_ERR_UNSUPPORTED), ((2894, 2894, 20), X.XRH_ERR_INVALID),
                       ((512, 512, 400), X.XRH_ERR_INVALID)):
        sz = np.array(size, np.int64)
        st = lib.xrh_volume_create(None, X._fp(vox), sz.ctypes.data_as(X._PI64), X._fp(i2p), X.XRH_F32,
                                   ctypes.byref(out))
        msg = lib.xrh_last_error().decode()
        assert st == want, (size, st, msg)
        if want == X.XRH_ERR_INVALID:
            assert "null context" in msg
