"""bench.py's command-line contract on the CPU: a GPU count other than 1 is
refused with a one-line JSON message before any device work."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    for n in (2, 8):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n), "--steps", "1"],
                           capture_output=True, text=True, timeout=120, cwd=ROOT)
        assert r.returncode == 2, r.stderr
        line = json.loads(r.stdout.strip().splitlines()[-1])
        assert line["n_gpus"] == n and "one GPU" in line["error"]
