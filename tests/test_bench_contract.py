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
    # Four sweeps: 70 state words + 18.5 geometry words + 48 B of ids per csu;
    # the single-pass compulsory bytes 18.5 words + 24 B.
    b = _bench()
    for w, total, comp in ((4, 402.0, 98.0), (8, 756.0, 172.0)):
        assert b.model_bytes(w) == 70 * w + 18.5 * w + 48 == total
        assert b.compulsory_bytes(w) == comp
    assert b.flux_bytes(4) == 274.0


def test_roofline_fractions_and_traffic_record(tmp_path, monkeypatch):
    # frac_model counts the algorithmic bytes, frac_measured the recorded
    # counter traffic, which applies only to the kernel source it was measured on
    b = _bench()
    sha = b.kernel_source_sha()
    rec = {"w-1000:fp32:stage_tile": dict(bytes_per_launch=160000, read_bytes=120000,
                                          write_bytes=40000, kernel_source_sha=sha),
           "w-2000:fp32:stage_tile": dict(bytes_per_launch=1, read_bytes=1, write_bytes=0,
                                          kernel_source_sha="0" * 16)}
    (tmp_path / "profiles").mkdir()
    (tmp_path / "profiles" / "traffic.json").write_text(json.dumps(rec))
    real_root = b.ROOT
    monkeypatch.setattr(b, "ROOT", str(tmp_path))
    monkeypatch.setattr(b, "kernel_source_sha", lambda: sha)
    r = b.roofline(1000, 0.001, "w-1000")   # 1000 cells in 1 us
    assert r["frac"] == r["frac_model"] == round(402.0 * 1000 / 1e-6 / 1e9 / 8000.0, 4)
    assert r["frac_compulsory"] == round(98.0 * 1000 / 1e-6 / 1e9 / 8000.0, 4)
    assert r["traffic"] == 160000 and r["frac_measured"] == round(160000 / 1e-6 / 1e9 / 8000.0, 4)
    assert r["measured_bytes_per_cell"] == 160.0
    stale = b.roofline(1000, 0.001, "w-2000")
    assert stale["traffic"] is None and stale["frac_measured"] is None
    assert "another kernel source" in stale["traffic_note"]
    assert b.roofline(1000, 0.001, "w-3000")["frac_measured"] is None
    monkeypatch.setattr(b, "ROOT", real_root)


def test_committed_traffic_record_is_well_formed():
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        rec = json.load(f)
    assert rec, "no measured traffic recorded"
    for key, e in rec.items():
        workload, prec, kernel = key.split(":")
        assert prec == "fp32" and kernel == "stage_tile"
        assert e["bytes_per_launch"] == e["read_bytes"] + e["write_bytes"] > 0
        assert len(e["kernel_source_sha"]) == 16 and e["command"]
