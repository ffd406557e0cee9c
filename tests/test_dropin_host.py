"""Host-only checks of the drop-in C++ classes' bookkeeping, built with AddressSanitizer and
UBSan and run on the CPU (no device calls): tests/cpp/raycaster_bookkeeping.cpp."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_raycaster_grow_after_allocate_stays_in_bounds(tmp_path):
    lib = ROOT / "xreg_amd" / "libxrh.so"
    if not lib.exists():
        pytest.skip("libxrh.so not built")
    exe = tmp_path / "raycaster_bookkeeping"
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           f"-I{ROOT / 'include'}", str(ROOT / "tests" / "cpp" / "raycaster_bookkeeping.cpp"),
           str(ROOT / "xreg_amd" / "cpp" / "xreg_hip.cpp"), f"-L{lib.parent}", "-lxrh",
           f"-Wl,-rpath,{lib.parent}", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0:protect_shadow_gap=0",
               UBSAN_OPTIONS="halt_on_error=1")
    res = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "raycaster bookkeeping: ok" in res.stdout
