"""Schedule autotuner for the minors kernels.

Picks, by measurement on the GPU, the kernel schedule of a GreenFKEngine:
one lane per (omega, l) job, with the 256-instance work unit split into 1, 2
or 4 blocks (smaller blocks free their SIMD sooner; larger ones restage the
frequency's table image less often), or the persistent substep-scheduled
kernels, the waves per SIMD those kernels are compiled for (1: no register
spills, 2: more latency hiding) and their work-unit size (instances of one
frequency handed to a block, at most 256).  Every candidate performs the same arithmetic in every lane, so the
choice never changes a result, only the time.

The sample is frequency-stratified: every k-th frequency of the grid with all
of its degrees, so the cost mix (low/high frequency, low/high degree) of the
full workload is preserved.  Results are cached in a JSON file keyed by the
workload signature and the device name.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

DEFAULT_CANDIDATES = (
    {"persistent": 0, "sph_split": 4},
    {"persistent": 0, "sph_split": 2},
    {"persistent": 0, "sph_split": 1},
    {"persistent": 1, "waves": 1, "unit": 128},
    {"persistent": 1, "waves": 1, "unit": 256},
    {"persistent": 1, "waves": 2, "unit": 256},
)
CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")


def _signature(engine, grid, eps, what):
    return (f"{what}|nk{engine.nem.nk}|nlay{engine.nem.nlay}|nnod{engine.nodes.nnod}|nf{grid.freqs.size}|"
            f"nkf{grid.nkf}|fmax{grid.freqs.max():.6g}|eps{eps:g}|{engine.variant}")


def tune(engine, grid, jre, eps, candidates=DEFAULT_CANDIDATES, stride=None, repeats=1, use_cache=True,
         what="minors", log=None):
    """Return (best_config, table of (config, kernel_ms)) and configure the engine with the best."""
    sig = _signature(engine, grid, eps, what)
    cache = {}
    if use_cache and os.path.exists(CACHE):
        try:
            with open(CACHE) as fh:
                cache = json.load(fh)
        except (OSError, ValueError):
            cache = {}
    if use_cache and sig in cache:
        best = cache[sig]["best"]
        engine.configure(**best)
        return best, cache[sig]["table"]
    nf = grid.freqs.size
    stride = stride or max(1, nf // 48)
    sel = np.arange(stride // 2, nf, stride)
    freqs, nwn = grid.freqs[sel], grid.nwn[sel]
    table = []
    for cand in candidates:
        engine.configure(**cand)
        best_ms = None
        for _ in range(repeats + 1):          # first run warms up allocations
            t0 = time.perf_counter()
            if what == "minors":
                engine.run_minors_resident(freqs, 0, nwn, grid.sigma, eps, jre, jre)
            else:
                engine.run_gfk_resident(freqs, 0, nwn, grid.sigma, eps, jre)
            wall = (time.perf_counter() - t0) * 1e3
            ms = float(engine.timing()[5]) or wall
            best_ms = ms if best_ms is None else min(best_ms, ms)
        table.append((dict(cand), best_ms))
        if log:
            log(f"autotune {cand}: {best_ms:.1f} ms on {int(nwn.sum())} solves")
    best = min(table, key=lambda t: t[1])[0]
    engine.configure(**best)
    if use_cache:
        cache[sig] = {"best": best, "table": table, "sample_freqs": int(sel.size)}
        try:
            with open(CACHE, "w") as fh:
                json.dump(cache, fh, indent=1)
        except OSError:
            pass
    return best, table
