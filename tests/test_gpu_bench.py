"""bench.py end to end on the GPU at a few steps: one JSON line with the contract's keys, a
positive throughput, the roofline, served, C3 and C4 blocks filled in."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def test_bench_json_line():
    print("case: bench.py --steps 2 (c3 registration, c4 storage leg, one served job, two mixed jobs)", flush=True)
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--steps", "2", "--warmup", "1", "--no-cpu",
                        "--plain-poses", "2", "--serve-jobs", "1", "--serve-gens", "2", "--c4-steps", "1",
                        "--serve-mix-jobs", "2", "--serve-mix-rate", "4"],
                       capture_output=True, text=True, timeout=400, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["value"] > 0
    rf = d["roofline"]
    assert rf["bound"] == "valu" and rf["kernel_ms"] > 0 and rf["algorithmic_gbs"] > 0
    if rf["frac"] is not None:
        assert 0 < rf["frac"] <= 1 and abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    if rf["hbm_frac"] is not None:
        assert 0 < rf["hbm_frac"] <= 1
    assert d["fp32_geometry_poses_per_s"] > 0
    assert d["served"]["jobs"] == 1 and d["c3_registration"]["levels"][0]["evals"] > 0
    assert d["autotune"]["variants_timed"] == 20
    assert d["dropin_poses_per_s"] > 0
    for st in ("f32", "f16", "bf16"):
        assert d["c4"][st]["poses_per_s"] > 0
    mix = d["served_mix"]
    assert mix["jobs"] == 2 and mix["same_work"]
    assert sum(c["jobs"] for c in mix["batched"]["classes"].values()) == 2
