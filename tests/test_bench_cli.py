"""bench.py's command line: host-side checks that need no GPU."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    # The build runs on one GPU: --gpus N > 1 exits with status 2 and a
    # message, before anything loads the library or touches a device.
    for n in (2, 8):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                            "--gpus", str(n)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 2, r.stderr
        assert f"--gpus {n} is not supported" in r.stderr
        assert r.stdout == ""
