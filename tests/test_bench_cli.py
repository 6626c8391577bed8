"""bench.py's command line on the CPU: it measures one GPU, so a request for
more exits with status 2 and says so before it loads the library."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "measures one GPU only (--gpus 2 requested)" in r.stderr
    assert r.stdout == ""  # no benchmark line
