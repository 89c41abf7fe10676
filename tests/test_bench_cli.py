"""bench.py's command line on the CPU: more than one GPU is refused with a
JSON message before anything touches a device."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    for n in (2, 8):
        proc = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                              capture_output=True, text=True, timeout=120)
        assert proc.returncode == 2
        line = json.loads(proc.stdout.strip().splitlines()[-1])
        assert "one GPU" in line["error"]
