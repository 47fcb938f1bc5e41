"""Periodic and open 3-D pair search of tensorsph.threed.neighbors3
(neighbors3.py:36-170): the HIP cell sort of libtsph, and its separate
O(N^2) kernel for the brute-force check."""
from __future__ import annotations

import torch

from ..neighbor_search import _search, cell_linked_pairs, periodic_pairs as _periodic_pairs


def box_pairs(positions: torch.Tensor, cutoff: float) -> torch.Tensor:
    """Unordered pairs of an open cloud (neighbors3.py:36-42)."""
    return cell_linked_pairs(positions, cutoff)


def _box_tensor(box, like: torch.Tensor) -> torch.Tensor:
    dim = like.shape[-1]
    out = torch.as_tensor(box, dtype=like.dtype, device=like.device).reshape(-1)
    if out.numel() == 1:
        out = out.expand(dim)
    if out.numel() != dim:
        raise ValueError(f"box must have 1 or {dim} edges, got {out.numel()}")
    return out.reshape(1, dim)


def minimum_image(delta: torch.Tensor, box) -> torch.Tensor:
    """delta - L round(delta / L), round half to even (neighbors3.py:45-53)."""
    box_t = _box_tensor(box, delta)
    return delta - box_t * torch.round(delta / box_t)


def periodic_pairs(positions: torch.Tensor, box, cutoff: float) -> torch.Tensor:
    """Unordered minimum-image pairs in a periodic box (neighbors3.py:71-142)."""
    return _periodic_pairs(positions, box, cutoff)


def brute_force_pairs_min_image(positions: torch.Tensor, box, cutoff: float,
                                chunk: int = 512) -> torch.Tensor:
    """The reference's O(N^2) minimum-image search (neighbors3.py:145-170):
    every pair i < j tested on the GPU, no cell machinery.  ``chunk`` is the
    reference's memory knob and has no effect here."""
    if positions.dim() != 2:
        raise ValueError("positions must be (N, D)")
    return _search(positions, cutoff, box=box, brute=True)
