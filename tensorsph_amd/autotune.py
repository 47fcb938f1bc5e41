"""Autotuner for the cell-tile shape of the resident 2-D stepper.

Times a few steps of the caller's own simulation for each candidate tile
shape (cells per workgroup tile along x and y) with the HIP-event sweep
timers of libtsph, keeps the fastest, and leaves the simulation configured
with it.  Changing the tile shape only re-orders the same particles, so the
trial steps advance the simulation like any other steps.
"""
from __future__ import annotations

import torch

DEFAULT_CANDIDATES = ((4, 4), (6, 4), (7, 4), (8, 4), (5, 6), (8, 6), (8, 8), (16, 4))


def tune_tiles(sim, candidates=DEFAULT_CANDIDATES, steps: int = 6, warm: int = 2) -> dict:
    """Returns {"best": (tx, ty), "ms": {shape: ms per step}}; sim ends on the best shape."""
    if sim.dim != 2:
        return {"best": None, "ms": {}}
    results = {}
    for shape in candidates:
        try:
            sim.configure(tiles=True, tile=shape)
        except ValueError:
            continue
        sim.step(warm)                      # includes the re-layout
        torch.cuda.synchronize()
        sim.set_timing(True)
        sim.step(steps)
        torch.cuda.synchronize()
        ms, n = sim.sweep_times()
        sim.set_timing(False)
        results[tuple(shape)] = sum(ms) / max(n, 1)
    best = min(results, key=results.get)
    sim.configure(tiles=True, tile=best)
    return {"best": best, "ms": {f"{k[0]}x{k[1]}": v for k, v in results.items()}}
