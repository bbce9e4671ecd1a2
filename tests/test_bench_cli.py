"""bench.py's contract on a host without a GPU: a run asking for more than one GPU exits at once
with status 2 and a message saying that this build measures one GPU (nothing touches a device)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 2
    assert "measures one GPU" in p.stderr and "--gpus 2" in p.stderr
    assert p.stdout == ""  # no JSON line
