"""bench.py's command-line contract on the host: more than one GPU is
refused with a message (this build runs on one GPU), for every workload,
before any device is touched."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("workload", ["single", "batch", "precision"])
def test_more_than_one_gpu_is_refused(workload):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2",
                        "--workload", workload], capture_output=True, text=True, timeout=120,
                       cwd=ROOT)
    assert r.returncode != 0
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]
