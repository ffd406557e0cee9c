"""bench.py's command line on the CPU: more than one GPU is refused with a message before any
device work, and the profile lookup only returns figures measured on the current kernel source."""
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


def test_profile_lookup_is_keyed_by_kernel_source(tmp_path):
    sys.path.insert(0, str(ROOT))
    import bench

    src = bench.kernel_source_key()
    assert len(src) == 12
    key = f"workload=cmaes-lambda100-384sq-patch-grad-ncc storage=f32 geometry=fp64 skip=1 smooth=5 src={src}"
    (tmp_path / "p.md").write_text(
        "# summary\n"
        f"counters: {key} valu_insts=2.5e9 vmem_insts=1.7e8 traffic_bytes=8.1e9 ta_busy_frac=0.86\n"
        f"counters: {key.replace(src, '0' * 12)} valu_insts=9e9\n")
    got = bench.profiled(key, tmp_path)
    assert got == {"valu_insts": 2.5e9, "vmem_insts": 1.7e8, "traffic_bytes": 8.1e9, "ta_busy_frac": 0.86}
    # a figure measured on other kernel code, or another configuration, is not reported
    assert bench.profiled(key.replace(src, "f" * 12), tmp_path) == {}
    assert bench.profiled(key.replace("384sq", "768sq"), tmp_path) == {}


def test_cpu_info_fields():
    sys.path.insert(0, str(ROOT))
    import bench

    info = bench.cpu_info()
    assert info["nproc"] >= 1 and 1 <= info["affinity_cpus"] <= info["nproc"]
    assert bench.device_cus() > 0
