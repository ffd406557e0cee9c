"""The float division / square root forms of the FP32 ray set-up round exactly like the host's
IEEE arithmetic (a single-ulp difference would flip step counts floor(chord/step) now and then)."""
import json
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parents[1] / "xreg_amd" / "_build" / "fp_rounding_check"


def test_setup_division_and_sqrt_are_correctly_rounded():
    print("case: fp_rounding_check 4M operand pairs", flush=True)
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print(res)
    assert res["sqrt_via_f64"] == 0
    assert res["div_via_f64"] == 0
