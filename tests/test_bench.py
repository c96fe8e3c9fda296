"""Host logic of bench.py (no GPU): the one-GPU contract and the roofline record."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_rejected():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_roofline_record():
    b = _bench()
    prof = {"smooth": {"total_ms": 10.0, "launches": 20, "bytes": 1.0e9}}
    r = b.roofline_of(prof, 256, "fp32")
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert abs(r["avg_launch_ms"] - 0.5) < 1e-12
    assert abs(r["achieved"] - 2000.0) < 1e-6          # 1 GB / 0.5 ms
    assert abs(r["frac"] - 0.25) < 1e-12
    assert r["traffic"] == b.SMOOTH256_TRAFFIC_BYTES["fp32"]
    assert "float" in r["kernel"]
    # the measured traffic is only known for the 256^3 workload it was taken on
    assert b.roofline_of(prof, 128, "fp64")["traffic"] is None
