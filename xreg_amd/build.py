"""Builds libxrh.so (the gfx950 HIP library behind include/xrh.h) in-tree.

Usage: python -m xreg_amd.build   (or xreg_amd.build.build())
Compiles every csrc/*.hip with hipcc for gfx950 and links xreg_amd/libxrh.so. Always
recompiles (parallel, one hipcc per source) so a stale library never survives a source edit.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libxrh.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no implicit FMA; the ray set-up must match the CPU reference's
#   separate multiply/add rounding (hot loops use explicit fmaf where intended).
# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU instructions.
COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


def _compile(src: Path, extra: list[str]) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, *extra, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False, extra: list[str] | None = None) -> Path:
    extra = list(extra or [])
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    with cf.ThreadPoolExecutor(max_workers=min(len(sources), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda s: _compile(s, extra), sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources")
    build_cpp(verbose)
    return LIB


CPP_LIB = HERE / "libxreg_hip.so"
DRIVER = BUILD / "dropin_driver"
FEATURES = BUILD / "dropin_features"
CXX = os.environ.get("CXX", "g++")
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def build_cpp(verbose: bool = False) -> Path:
    """The drop-in C++ classes (plain g++, linked against libxrh.so only through the C-ABI) and
    the driver that exercises them like Intensity2D3DRegi."""
    src = HERE / "cpp" / "xreg_hip.cpp"
    tmp = CPP_LIB.with_suffix(".so.tmp")
    cmd = [CXX, *CXXFLAGS, "-shared", str(src), f"-L{HERE}", "-lxrh", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed for xreg_hip.cpp:\n{res.stderr}")
    os.replace(tmp, CPP_LIB)
    for name, exe in (("dropin_driver.cpp", DRIVER), ("dropin_features.cpp", FEATURES)):
        drv = HERE.parent / "tests" / "cpp" / name
        cmd = [CXX, *CXXFLAGS, str(drv), f"-L{HERE}", "-lxreg_hip", "-lxrh", "-Wl,-rpath,$ORIGIN/..", "-o", str(exe)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"g++ failed for {name}:\n{res.stderr}")
    fp = HERE.parent / "tests" / "cpp" / "fp_rounding_check.hip"
    cmd = [HIPCC, *COMMON, str(fp), "-o", str(FP_CHECK)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for fp_rounding_check.hip:\n{res.stderr}")
    if verbose:
        print(f"built {CPP_LIB}, {DRIVER}, {FEATURES} and {FP_CHECK}")
    return CPP_LIB


FP_CHECK = BUILD / "fp_rounding_check"


ASAN_DIR = BUILD / "asan"


def build_asan(verbose: bool = False) -> Path:
    """Host-side AddressSanitizer + UBSan build of libxrh (device code is built as usual) and of
    the drop-in drivers with the C++ classes compiled in, all with ROCm's clang so one sanitizer
    runtime serves the process. Run the drop-in tests against them with
    XREG_AMD_DROPIN_DIR=<ASAN_DIR> and
    ASAN_OPTIONS=detect_leaks=0:verify_asan_link_order=0:protect_shadow_gap=0."""
    ASAN_DIR.mkdir(parents=True, exist_ok=True)
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-omit-frame-pointer", "-g"]
    flags = [f for f in COMMON if f != "-O3"] + ["-O1"]
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        futs = []
        for src in sorted(CSRC.glob("*.hip")):
            obj = ASAN_DIR / (src.stem + ".o")
            futs.append(ex.submit(subprocess.run, [HIPCC, *flags, *san, "-c", str(src), "-o", str(obj)],
                                  capture_output=True, text=True))
            objs.append(obj)
        for f, src in zip(futs, sorted(CSRC.glob("*.hip"))):
            res = f.result()
            if res.returncode != 0:
                raise RuntimeError(f"hipcc (asan) failed for {src.name}:\n{res.stderr}")
    lib = ASAN_DIR / "libxrh.so"
    res = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(lib), *map(str, objs)],
                         capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link (asan) failed:\n{res.stderr}")
    clang = Path(HIPCC).resolve().parent.parent / "llvm" / "bin" / "clang++"
    root = HERE.parent
    for name in ("dropin_driver", "dropin_features"):
        cmd = [str(clang), "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
               f"-I{root / 'include'}", str(root / "tests" / "cpp" / f"{name}.cpp"), str(HERE / "cpp" / "xreg_hip.cpp"),
               f"-L{ASAN_DIR}", "-lxrh", "-Wl,-rpath,$ORIGIN", "-o", str(ASAN_DIR / name)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"clang++ (asan) failed for {name}:\n{res.stderr}")
    for o in objs:
        o.unlink()
    if verbose:
        print(f"built the drop-in drivers in {ASAN_DIR}")
    return ASAN_DIR


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)
    if LIB.exists():
        LIB.unlink()


if __name__ == "__main__":
    if sys.argv[1:] == ["--asan"]:
        build_asan(verbose=True)
    else:
        build(verbose=True, extra=sys.argv[1:])
