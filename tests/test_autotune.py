"""The launch-shape autotuner: its cache (CPU) and a short tuning run (GPU)."""
import json

import pytest

from gpueularsolver2d_amd import autotune, naca0012_mesh


def test_cache_lookup_and_best_config_hit(tmp_path, monkeypatch):
    # a cached entry for this device/scheme/size bucket is returned without tuning
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "cache.json"))
    dev = autotune._device_name()
    key = autotune._key("fp32", 2, True, 200000, dev)
    with open(autotune.CACHE, "w") as f:
        json.dump({key: dict(tile_cells=259, threads=512, ms_per_step=0.05, ncell=198578)}, f)
    assert autotune.lookup("fp32", 2, True, 200000) == dict(tile_cells=259, threads=512,
                                                             ms_per_step=0.05, ncell=198578)
    # sizes in the same half-octave bucket share the entry; other schemes do not
    assert autotune.lookup("fp32", 2, True, 210000) is not None
    assert autotune.lookup("fp64", 2, True, 200000) is None
    assert autotune.lookup("fp32", 1, True, 200000) is None
    hit_mesh = type("M", (), {"ncell": 205000})()  # only its size is read on a hit
    assert autotune.best_config(hit_mesh, "fp32", 2, True) == dict(tile_cells=259, threads=512)


@pytest.mark.gpu
def test_tune_times_candidates_and_skips_invalid(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "cache.json"))
    m = naca0012_mesh(6000, seed=12)
    cands = [(0, 256), (64, 128), (6000, 512)]     # the last tile needs too much LDS
    r = autotune.tune(m, precision="fp32", candidates=cands, steps=10, write_cache=True)
    rows = r["table"]
    assert len(rows) == 3 and "error" in rows[2] and "error" not in rows[0]
    ok = [row for row in rows if "error" not in row]
    assert all(row["ms_per_step"] > 0 and row["kernel_ms"] > 0 for row in ok)
    assert r["ms_per_step"] == min(row["ms_per_step"] for row in ok)
    assert (r["tile_cells"], r["threads"]) in [(row["tile_cells"], row["threads"]) for row in ok]
    # the result is cached under this device and size bucket
    assert autotune.lookup("fp32", 2, True, m.ncell)["tile_cells"] == r["tile_cells"]


@pytest.mark.gpu
def test_tune_batch_picks_from_timed_candidates():
    # a batch of uneven cases: the default candidates around the library's
    # own choice, each timed on the whole batch
    meshes = [naca0012_mesh(3000, seed=11), naca0012_mesh(8000, seed=12)]
    flows = [dict(mach=0.5, alpha_deg=1.0, cfl=1.5), dict(mach=0.8, alpha_deg=1.25, cfl=1.5)]
    r = autotune.tune_batch(meshes, [0, 1, 1, 0], [flows[0], flows[1], flows[0], flows[1]],
                            precision="fp32", steps=6)
    rows = [row for row in r["table"] if "error" not in row]
    assert len(rows) >= 5 and rows[0]["requested_tile_cells"] == 0 and rows[0]["threads"] == 512
    assert all(row["ms_per_step"] > 0 and row["kernel_ms"] > 0 for row in rows)
    assert r["ms_per_step"] == min(row["ms_per_step"] for row in rows)
    assert (r["tile_cells"], r["threads"]) in [(row["tile_cells"], row["threads"]) for row in rows]
