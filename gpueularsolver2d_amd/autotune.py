"""Autotuner of the tiled kernels' launch shape.

The tile kernels run one workgroup per tile of cells.  Three things follow from the tile
size:
  * the LDS footprint, and with the registers, how many workgroups share a CU;
  * the ring overhead (cells next to a tile are recomputed by each tile that needs them);
  * how the tiles fill the device's workgroup slots.

The workgroup size sets how many passes each phase needs over the tile.  With tile_cells = 0
the library sizes the tiles so that they fill whole waves of slots.  tune_batch() times that
choice at several workgroup sizes, beside explicit (tile_cells, threads) candidates, with
short runs of the real batch (many cases on uneven meshes in every launch), and keeps the
fastest per time step.  tune() does the same for one mesh and caches the result in
autotune_cache.json next to this file, keyed by the device name, precision, scheme and mesh
size bucket.

Two knobs of a colouring design do not exist here: cases per workgroup is one tile of one
case (a tile's LDS neighbourhood belongs to one mesh; a small case simply has fewer tiles),
and there is no edge colouring (each edge writes its own LDS slot and each cell gathers its
three slots in a fixed order, which keeps the FP64 sums in the reference's order).
"""
from __future__ import annotations

import json
import math
import os

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


def default_batch_candidates(auto_tile):
    """Launch shapes to time for a batch: the library's own tile choice at
    three workgroup sizes, and tiles of 0.8, 0.9, 1.1 and 1.2 times its
    512-thread choice at 512 threads."""
    c = [(0, 256), (0, 384), (0, 512)]
    c += [(max(64, int(round(auto_tile * f))), 512) for f in (0.8, 0.9, 1.1, 1.2)]
    return c


def tune_batch(meshes, case_mesh, flows, precision="fp32", order=2, smoothing=True,
               candidates=None, steps=20, verbose=False):
    """Times each candidate (tile_cells, threads) launch shape on a batch of
    cases (fvm_batch_*: every launch covers every case) and returns the
    fastest as {tile_cells, threads, ms_per_step, table}.  Each candidate runs
    10 steps, then `steps` timed steps from the device events around the run
    (fvm_batch_get_timing); its stage kernel is also timed alone.
    candidates: None picks default_batch_candidates around the library's
    choice for 512 threads."""
    from .solver import Batch, FvmError
    table = []

    def measure(tile, threads):
        try:
            b = Batch(meshes, case_mesh, flows, precision=precision, order=order,
                      smoothing=smoothing, tile_cells=tile, threads=threads)
            b.run(10, check=False)
            b.run(10 + steps, check=False)
            ms, csu = b.timing()
            kt = b.kernel_times(reps=5)
            cfg = b.config()
            row = dict(tile_cells=cfg["tile_cells"], threads=cfg["threads"],
                       requested_tile_cells=tile, ms_per_step=ms / steps,
                       mcsu_per_s=round(csu / ms / 1e3, 1) if ms > 0 else None,
                       kernel_ms=kt[0], kernel2_ms=kt[1], tiles=cfg["tiles"],
                       blocks_per_cu=cfg["blocks_per_cu"])
            b.close()
        except FvmError as e:  # e.g. the tile's LDS footprint exceeds the device limit
            row = dict(tile_cells=tile, threads=threads, error=str(e))
        table.append(row)
        if verbose:
            print(row, flush=True)
        return row

    if candidates is None:
        first = measure(0, 512)
        auto = first.get("tile_cells", 256)
        candidates = [c for c in default_batch_candidates(auto) if c != (0, 512)]
    for tile, threads in candidates:
        measure(tile, threads)
    ok = [r for r in table if "error" not in r]
    best = min(ok, key=lambda r: r["ms_per_step"])
    return dict(tile_cells=best["tile_cells"], threads=best["threads"],
                ms_per_step=best["ms_per_step"], table=table)


def tune(mesh, precision="fp32", order=2, smoothing=True, candidates=None, steps=60, flow=None,
         write_cache=True, verbose=False):
    """Times each candidate launch shape on one case of `mesh` (tune_batch
    with one case); returns the best as {tile_cells, threads, ms_per_step,
    ncell, table} and caches it per device, scheme and size bucket."""
    flow = flow or dict(mach=0.8, alpha_deg=1.25, cfl=1.5)
    r = tune_batch([mesh], [0], [flow], precision, order, smoothing,
                   candidates or DEFAULT_CANDIDATES, steps, verbose)
    out = dict(tile_cells=r["tile_cells"], threads=r["threads"], ms_per_step=r["ms_per_step"],
               ncell=mesh.ncell, table=r["table"])
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
