"""GPU linked-cell pair search with the reference's contract.

``cell_linked_pairs`` (neighbor_search.py:44-120), ``periodic_pairs``
(threed/neighbors3.py:71-142), ``brute_force_pairs`` (:29-41) and
``PairListCache`` (:166-227): same signatures, same validation and
exceptions, same output -- unordered pairs ``(i, j)``, ``i < j``, sorted
lexicographically, int64, on the input device.  The search itself is the
HIP cell sort of libtsph (tsph_pairs_search / tsph_pairs_export); the
brute-force searches are libtsph's separate O(N^2) kernel
(tsph_pairs_brute), so the two can check each other as the reference's
tests do.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from . import _lib


def _validate(positions, cutoff) -> tuple[int, int]:
    if not isinstance(positions, torch.Tensor):
        raise TypeError("positions must be a torch.Tensor")
    if positions.ndim != 2:
        raise ValueError("positions must have shape (particle_count, dimension)")
    if positions.shape[1] < 1:
        raise ValueError("positions must have at least one spatial dimension")
    if not positions.is_floating_point():
        raise TypeError("positions must use a floating-point dtype")
    if not math.isfinite(float(cutoff)) or cutoff <= 0:
        raise ValueError("cutoff must be finite and positive")
    if positions.numel() and not bool(torch.isfinite(positions).all()):
        raise ValueError("positions must contain only finite values")
    return int(positions.shape[0]), int(positions.shape[1])


def _empty(device) -> torch.Tensor:
    return torch.empty((0, 2), dtype=torch.long, device=device)


def _dtype_code(t: torch.Tensor) -> tuple[torch.Tensor, int]:
    if t.dtype == torch.float64:
        return t, _lib.F64
    if t.dtype == torch.float32:
        return t, _lib.F32
    # half/bfloat16: search in float32 (the reference would search in the
    # narrow type; those dtypes are not used on this path)
    return t.to(torch.float32), _lib.F32


def _search(positions, cutoff, box=None, fluid_count=None, with_cells=False, brute=False):
    n, dim = _validate(positions, cutoff)
    if dim > 3:
        raise ValueError("positions must have at most 3 dimensions on the GPU search")
    dev = _lib.require_gpu(positions)
    if n < 2:
        return (_empty(dev), None) if with_cells else _empty(dev)
    pos, code = _dtype_code(positions.detach().contiguous())
    ctx = _lib.default_context(dev)
    stream = _lib.stream_handle(dev)
    count = C.c_int64(0)
    box_arr = None
    if box is not None:
        edges = torch.as_tensor(box, dtype=torch.float64).reshape(-1).tolist()
        if len(edges) == 1:
            edges = edges * dim
        if len(edges) != dim:
            raise ValueError(f"box must have 1 or {dim} edges, got {len(edges)}")
        if any(e <= 0.0 for e in edges):
            raise ValueError("box edges must be positive")
        box_arr = (C.c_double * 3)(*(edges + [0.0] * (3 - dim)))
    fc = -1 if fluid_count is None else int(fluid_count)
    L = _lib.lib()
    if brute:
        _lib.check(L.tsph_pairs_brute(ctx.handle, _lib.ptr(pos), code, n, dim, float(cutoff),
                                      box_arr, stream, C.byref(count)))
    else:
        _lib.check(L.tsph_pairs_search(ctx.handle, _lib.ptr(pos), code, n, dim, float(cutoff),
                                       box_arr, fc, stream, C.byref(count)))
    out = torch.empty((count.value, 2), dtype=torch.long, device=dev)
    if count.value:
        _lib.check(L.tsph_pairs_export(ctx.handle, _lib.ptr(out), stream))
    if not with_cells:
        return out
    cells = torch.empty((n, dim), dtype=torch.long, device=dev)
    keys = torch.empty(n, dtype=torch.long, device=dev)
    order = torch.empty(n, dtype=torch.long, device=dev)
    ncells = C.c_int64(0)
    _lib.check(L.tsph_cells_export(ctx.handle, _lib.ptr(cells), _lib.ptr(keys), _lib.ptr(order),
                                   C.byref(ncells), stream))
    return out, {"cells": cells, "keys": keys, "order": order, "ncells": ncells.value}


def cell_linked_pairs(positions: torch.Tensor, cutoff: float) -> torch.Tensor:
    """Exact unordered pairs within ``cutoff`` (pairs at the cutoff included)."""
    return _search(positions, cutoff)


def brute_force_pairs(positions: torch.Tensor, cutoff: float) -> torch.Tensor:
    """The reference's O(N^2) search (neighbor_search.py:29-41): every pair
    i < j tested on the GPU, with no cell machinery (at most 2^20 particles)."""
    return _search(positions, cutoff, brute=True)


def cell_linked_pairs_with_cells(positions: torch.Tensor, cutoff: float):
    """Pairs plus the cell bookkeeping (cells - min, row-major keys, stable order)."""
    return _search(positions, cutoff, with_cells=True)


def periodic_pairs(positions: torch.Tensor, box, cutoff: float) -> torch.Tensor:
    """Minimum-image pairs in a periodic rectangular box (neighbors3.py:71-142)."""
    if positions.dim() != 2:
        raise ValueError("positions must be (N, D)")
    return _search(positions, cutoff, box=box)


def box_pairs(positions: torch.Tensor, cutoff: float) -> torch.Tensor:
    return cell_linked_pairs(positions, cutoff)


def drop_wall_wall_pairs(pairs: torch.Tensor, fluid_count: int) -> torch.Tensor:
    """droplet_impact.drop_wall_wall_pairs (droplet_impact.py:791-807)."""
    if pairs.numel() == 0:
        return pairs
    i, j = pairs.unbind(dim=1)
    usable = (i < fluid_count) | (j < fluid_count)
    if bool(usable.all()):
        return pairs
    return pairs[usable]


def max_displacement(positions: torch.Tensor, anchor: torch.Tensor) -> float:
    dev = _lib.require_gpu(positions, anchor)
    pos, code = _dtype_code(positions.detach().contiguous())
    anc = anchor.detach().to(pos.dtype).contiguous()
    out = torch.zeros((), dtype=pos.dtype, device=dev)
    _lib.check(_lib.lib().tsph_max_displacement(_lib.ptr(pos), _lib.ptr(anc), code, pos.shape[0],
                                                pos.shape[1], _lib.ptr(out), _lib.stream_handle(dev)))
    return float(out)


class PairListCache:
    """Verlet-skin pair cache (neighbor_search.py:166-227).

    Builds at ``cutoff + skin`` and rebuilds once any particle has moved more
    than ``skin / 2`` since the last build; ``fluid_count`` prunes pairs with
    both ends at or beyond it when the list is built.
    """

    def __init__(self, cutoff: float, skin: float, fluid_count: int | None = None):
        if not math.isfinite(float(cutoff)) or cutoff <= 0:
            raise ValueError("cutoff must be finite and positive")
        if not math.isfinite(float(skin)) or skin < 0:
            raise ValueError("skin must be finite and non-negative")
        if fluid_count is not None and fluid_count < 0:
            raise ValueError("fluid_count must be non-negative")
        self.cutoff = float(cutoff)
        self.skin = float(skin)
        self.fluid_count = fluid_count
        self._pairs = None
        self._anchor = None
        self.rebuilds = 0
        self.reuses = 0
        self.pruned = 0

    def displacement(self, positions: torch.Tensor) -> float:
        if self._anchor is None or self._anchor.shape != positions.shape:
            return float("inf")
        return max_displacement(positions, self._anchor)

    def _build(self, positions: torch.Tensor) -> torch.Tensor:
        pairs = cell_linked_pairs(positions, self.cutoff + self.skin)
        if self.fluid_count is not None and pairs.numel():
            i, j = pairs.unbind(dim=1)
            keep = (i < self.fluid_count) | (j < self.fluid_count)
            self.pruned += int(keep.numel() - keep.sum())
            pairs = pairs[keep]
        return pairs

    def get(self, positions: torch.Tensor):
        if self._pairs is None or self.displacement(positions) > 0.5 * self.skin:
            self._pairs = self._build(positions)
            self._anchor = positions.detach().clone()
            self.rebuilds += 1
            return self._pairs, True
        self.reuses += 1
        return self._pairs, False
