"""Autotuner of the tiled kernels' launch shape.

The tile kernels run one workgroup per tile of cells.  Three things follow from the tile
size:
  * the LDS footprint, and with the registers, how many workgroups share a CU;
  * the ring overhead (cells next to a tile are recomputed by each tile that needs them);
  * how the tiles fill the device's workgroup slots.

The workgroup size sets how many passes each phase needs over the tile.  With tile_cells = 0
the library sizes the tiles so that they fill whole waves of slots.  tune() times that choice
at several workgroup sizes, beside explicit (tile_cells, threads) candidates, with short runs
on the real mesh, and keeps the fastest per time step.  Results are cached in
autotune_cache.json next to this file.  The cache key is the device name, precision, scheme
and mesh size bucket.
"""
from __future__ import annotations

import json
import math
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "autotune_cache.json")

DEFAULT_CANDIDATES = [(0, 128), (0, 192), (0, 256), (0, 384), (0, 512)] + [
    (t, n) for t in (96, 128, 160, 192, 256) for n in (128, 256)]


def _key(precision, order, smoothing, ncell, device_name):
    bucket = int(round(math.log2(max(ncell, 1)) * 2))  # half-octave size buckets
    return f"{device_name}|{precision}|o{order}|s{int(bool(smoothing))}|b{bucket}"


def _device_name():
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.get_device_name(0)
    except Exception:
        pass
    return "unknown"


def load_cache():
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            return json.load(f)
    return {}


def lookup(precision, order, smoothing, ncell):
    return load_cache().get(_key(precision, order, smoothing, ncell, _device_name()))


def tune(mesh, precision="fp32", order=2, smoothing=True, candidates=None, steps=60, flow=None,
         write_cache=True, verbose=False):
    """Times each candidate launch shape on `mesh`; returns the best as a dict
    {tile_cells, threads, ms_per_step, table}."""
    from .solver import FvmError, Solver
    flow = flow or dict(mach=0.8, alpha_deg=1.25, cfl=1.5)
    candidates = candidates or DEFAULT_CANDIDATES
    table = []
    for tile, threads in candidates:
        try:
            s = Solver(mesh, precision=precision, order=order, smoothing=smoothing, tile_cells=tile,
                       threads=threads, **flow)
            s.run(10)
            t0 = time.perf_counter()
            s.run(10 + steps)
            wall = (time.perf_counter() - t0) / steps * 1e3
            ms, _ = s.timing()
            kt = s.kernel_times(reps=20)
            cfg = s.config()
            row = dict(tile_cells=cfg["tile_cells"], threads=threads, requested_tile_cells=tile,
                       ms_per_step=ms / steps, wall_ms_per_step=wall, flux_tile_ms=kt[0],
                       update_tile_ms=kt[1], tiles=cfg["tiles"], blocks_per_cu=cfg["blocks_per_cu"])
            s.close()
        except FvmError as e:  # e.g. the tile's LDS footprint exceeds the device limit
            row = dict(tile_cells=tile, threads=threads, error=str(e))
        table.append(row)
        if verbose:
            print(row, flush=True)
    ok = [r for r in table if "error" not in r]
    best = min(ok, key=lambda r: r["ms_per_step"])
    out = dict(tile_cells=best["tile_cells"], threads=best["threads"],
               ms_per_step=best["ms_per_step"], ncell=mesh.ncell, table=table)
    if write_cache:
        cache = load_cache()
        cache[_key(precision, order, smoothing, mesh.ncell, _device_name())] = {
            k: out[k] for k in ("tile_cells", "threads", "ms_per_step", "ncell")}
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return out


def best_config(mesh, precision="fp32", order=2, smoothing=True, retune=False):
    """Cached launch shape for this device/scheme/size, tuning on a miss."""
    hit = None if retune else lookup(precision, order, smoothing, mesh.ncell)
    if hit:
        return dict(tile_cells=hit["tile_cells"], threads=hit["threads"])
    r = tune(mesh, precision, order, smoothing)
    return dict(tile_cells=r["tile_cells"], threads=r["threads"])
