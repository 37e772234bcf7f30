"""bench.py's command-line contract, checked on the CPU: one GPU only, so a
request for more exits with a message instead of a result line."""
import os
import subprocess
import sys

from conftest import ROOT


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert not r.stdout.strip().startswith("{")  # no JSON result line
