"""bench.py's command line on the CPU: it parses, and a multi-GPU request exits with the
one-GPU message instead of running."""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, r.stderr
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_bench_help_lists_the_workloads():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for w in ("binary3d", "potts3d", "batched", "hbm"):
        assert w in r.stdout
