"""Autotuner for the sweep layout of the resident 2-D stepper.

Times a few steps of the caller's own simulation for each candidate layout
with the HIP-event sweep timers of libtsph, keeps the fastest, and leaves the
simulation configured with it.  A candidate is a cell-tile shape (cells per
workgroup tile along x and y) for the LDS force sweep, "lists" for the
force sweep over the per-row slot lists (the density and wall sweeps run over
those lists in every layout), or "halo" for the slot-list density and force
sweeps evaluated from each workgroup's neighbourhood rows staged in LDS.
Changing the layout only re-orders the same particles, so the trial steps
advance the simulation like any other steps.
"""
from __future__ import annotations

import torch

DEFAULT_CANDIDATES = ((4, 4), (6, 4), (7, 4), (8, 4), (5, 6), (8, 6), (8, 8), (16, 4), "lists", "halo")
_NAMED = ("lists", "halo")


def tune_tiles(sim, candidates=DEFAULT_CANDIDATES, steps: int = 6, warm: int = 2) -> dict:
    """Returns {"best": (tx, ty), "lists" or "halo", "ms": {layout: ms per step}}; sim ends on
    the best."""
    if sim.dim != 2:
        return {"best": None, "ms": {}}
    results = {}
    def use(shape):
        if shape == "lists":
            sim.configure(tiles=False)
        elif shape == "halo":
            sim.configure(tiles=False, halo=True)
        else:
            sim.configure(tiles=True, tile=shape)

    for shape in candidates:
        try:
            use(shape)
        except ValueError:
            continue
        sim.step(warm)                      # includes the re-layout
        torch.cuda.synchronize()
        sim.set_timing(True)
        sim.step(steps)
        torch.cuda.synchronize()
        ms, n = sim.sweep_times()
        sim.set_timing(False)
        results[shape if shape in _NAMED else tuple(shape)] = sum(ms) / max(n, 1)
    best = min(results, key=results.get)
    use(best)
    name = {k: (k if k in _NAMED else f"{k[0]}x{k[1]}") for k in results}
    return {"best": best, "ms": {name[k]: v for k, v in results.items()}}
