"""Host logic of bench.py (no GPU): the one-GPU contract and the roofline record."""
import importlib.util
import os
import subprocess
import sys

import numpy as np

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


def test_timed_iterations_cover_the_post_training_window():
    """The TO run is always 80 iterations with 15 training iterations; K timed steps are
    spread over iterations 15..79 whatever K is, so the headline does not depend on the
    window the flags would otherwise pick."""
    b = _bench()

    class A:
        pass
    a = A()
    a.steps, a.warmup = 65, 15
    assert b.timed_iterations(a) == list(range(15, 80))
    a.steps, a.warmup = 20, 5
    t = b.timed_iterations(a)
    assert len(t) == 20 and t[0] == 15 and t[-1] == 79 and t == sorted(set(t))
    assert max(np.diff(t)) - min(np.diff(t)) <= 1
    a.steps, a.warmup = 500, 0
    assert b.timed_iterations(a) == list(range(15, 80))
    a.steps, a.warmup = 10, 30                # more warmup than training: the window starts later
    t = b.timed_iterations(a)
    assert t[0] == 30 and t[-1] == 79 and len(t) == 10
