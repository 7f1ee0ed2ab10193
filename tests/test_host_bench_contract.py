"""bench.py's command-line contract on a machine without a GPU: one GPU only (--gpus N>1
exits with a message naming the reason)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300, cwd=ROOT)
    assert r.returncode != 0
    msg = (r.stdout + r.stderr).lower()
    assert "one gpu" in msg or "single gpu" in msg or "1 gpu" in msg, msg[-500:]
