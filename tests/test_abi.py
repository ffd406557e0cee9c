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
    assert ctypes.sizeof(X.SimParams) == 16
    assert X.View.cam.offset == 16
