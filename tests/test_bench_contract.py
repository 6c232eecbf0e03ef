"""The parts of bench.py's output contract that need no GPU: a run asking for
more than one GPU is refused with a message before anything touches a device,
and the HBM traffic per Q it reports comes from the committed rocprofv3 byte
counter summary."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench_module():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout.strip() == ""


def test_traffic_comes_from_the_committed_summary():
    bench = _bench_module()
    alg = 24 + 6 * 4 + 6 * 6 * 8  # config 4 FP32: Q in, values and vectors out
    for path in ("learned+binned fp32", "binned fp32"):
        v = bench.measured_traffic(4, path)
        assert v is not None and v > alg, (path, v)
    assert bench.measured_traffic(4, "no such path") is None
    assert bench.measured_traffic(99, "binned fp32") is None
