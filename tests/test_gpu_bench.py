"""bench.py end to end on the GPU at a few steps: one JSON line with the contract's keys, a
positive throughput, the roofline and served blocks filled in."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def test_bench_json_line():
    print("case: bench.py --steps 2 (c3 registration and one served job)", flush=True)
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--steps", "2", "--warmup", "1", "--no-cpu",
                        "--plain-poses", "2", "--serve-jobs", "1", "--serve-gens", "2", "--c3-regi"],
                       capture_output=True, text=True, timeout=240, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["value"] > 0
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["peak"] == 8000.0 and rf["achieved"] > 0
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    assert d["served"]["jobs"] == 1 and d["c3_registration"]["levels"][0]["evals"] > 0
    assert d["autotune"]["variants_timed"] == 20
    assert d["dropin_poses_per_s"] > 0
