"""bench.py's host-side pieces that need no GPU: the committed HBM-traffic
summary parses, covers the forward sweep and every kernel of an iteration,
and the --gpus contract refuses more than one GPU."""
import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_traffic_summary_parses(bench):
    fwd = bench.measured_traffic(bench.TRAFFIC_SIZE, "k_sweep_tile<true>")
    it = bench.measured_traffic(bench.TRAFFIC_SIZE)
    assert fwd is not None and fwd > 0
    # the iteration total includes kernels whose names hold spaces
    face = bench.measured_traffic(bench.TRAFFIC_SIZE, "k_face<0, false, double>")
    assert face is not None and face > 0
    assert it > fwd + face
    # a different workload has no summary
    assert bench.measured_traffic("64x32x32") is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
