"""bench.py's command line, without a GPU: more than one GPU is refused
with exit status 2 before anything touches a device, and the secondary
workloads can be run one at a time."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_legs_listed():
    r = _bench("--help")
    assert r.returncode == 0
    for leg in ("batch", "latency", "slab", "fcc_f32", "autograd", "lj", "alloc"):
        assert leg in r.stdout
