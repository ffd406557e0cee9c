"""bench.py's command line on the CPU: more than one GPU is refused with a message before any
device work, and the profile traffic lookup finds the committed summaries."""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_profile_traffic_lookup():
    sys.path.insert(0, str(ROOT))
    import bench

    key = "workload=cmaes-lambda100-384sq-patch-grad-ncc storage=f32 geometry=fp64 skip=1 smooth=5"
    assert bench.profiled_traffic(key) > 1e9
    assert bench.profiled_traffic(key.replace("384sq", "768sq")) > 1e10
    assert bench.profiled_traffic("workload=none") is None
