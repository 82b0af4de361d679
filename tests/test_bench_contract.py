"""bench.py's command-line contract, checked on the CPU: more than one GPU
is refused with a message (this build measures one GPU), and every
workload the measurement notes name is offered."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=300, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2, r.stderr[-2000:]
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_workloads_offered():
    r = _bench("--help")
    assert r.returncode == 0
    for w in ("c3", "c1", "c2", "c4", "serve", "union", "accuracy", "ins", "hbm", "dropin"):
        assert w in r.stdout
