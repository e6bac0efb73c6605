"""bench.py's contract pieces that run without a GPU: the one-GPU refusal, the
byte model behind roofline.achieved, and the measured-traffic lookup."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""  # no JSON line for a configuration that was not measured


def test_byte_model_totals():
    # Four sweeps: 70 state words + 18.5 geometry words + 48 B of ids per csu.
    b = _bench()
    for w, total in ((4, 402.0), (8, 756.0)):
        assert b.flux_bytes(w) + b.update_bytes(w) == 70 * w + 18.5 * w + 48 == total
    # The two-kernel split: sweeps 1+2 (gradient, flux) and 3+4 (smoothing, update).
    assert b.flux_bytes(4) == 274.0 and b.update_bytes(4) == 128.0


def test_traffic_lookup_matches_the_recorded_counters():
    b = _bench()
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        rec = json.load(f)
    assert rec, "no measured traffic recorded"
    for key, val in rec.items():
        workload, prec, kernel = key.split(":")
        assert b.traffic_from_profiles(workload, prec, kernel) == val
        assert val > 0
    assert b.traffic_from_profiles("naca0012-1", "fp32", "stage_tile") is None
