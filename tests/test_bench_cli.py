"""bench.py's command line on the CPU: the path runs on one GPU, so asking
for more exits with status 2 and a JSON message before any GPU work; the
options the workloads use are all declared."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_help_lists_every_workload_option():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--config", "--batch", "--yt", "--cube", "--host-buffers", "--copy-probe",
                "--no-cpu", "--no-fp64"):
        assert opt in r.stdout
