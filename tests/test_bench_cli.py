"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with a message and a non-zero status (the solver is not sharded)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode != 0
    doc = json.loads(out.stdout.strip().splitlines()[-1])
    assert doc["n_gpus"] == 2 and "one GPU only" in doc["error"]
