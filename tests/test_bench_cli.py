"""bench.py's command line, without a GPU: more than one GPU is refused with
a JSON message, and the help lists the workload selectors."""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _bench(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    res = _bench("--gpus", "2")
    assert res.returncode != 0
    msg = json.loads(res.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_help_lists_the_workloads():
    res = _bench("--help")
    assert res.returncode == 0
    for word in ("--only", "--schedule", "--per-body-sample", "--lambdas", "--copy-calib"):
        assert word in res.stdout
