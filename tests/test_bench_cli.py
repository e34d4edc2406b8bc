"""bench.py's command-line contract on the CPU: more than one GPU is refused
with a message (this build measures one GPU) before anything touches a
device."""

import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2, (r.returncode, r.stderr[-500:])
    line = r.stdout.strip().splitlines()[-1]
    assert "one GPU" in json.loads(line)["error"]
