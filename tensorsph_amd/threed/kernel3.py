"""3-D smoothing kernels of tensorsph.threed.kernel3 (kernel3.py:48-215).

Kernel values are computed by libtsph (tsph_kernel_eval, kind 1 = 3-D cubic
spline with its radial Laplacian); the lattice sums used to calibrate the
particle mass are host-side case set-up, as in the reference.
"""
from __future__ import annotations

import math

import torch

from .. import _lib
from ..cases import lattice_kernel_sum_3d

SUPPORT_FACTOR = 2.0


def sigma3(h: float) -> float:
    """Cubic-spline normalisation 1 / (pi h^3) (kernel3.py:48-50)."""
    return 1.0 / (math.pi * h ** 3)


def _code(t: torch.Tensor) -> int:
    if t.dtype == torch.float64:
        return _lib.F64
    if t.dtype == torch.float32:
        return _lib.F32
    raise TypeError(f"kernel arguments must be float32 or float64, got {t.dtype}")


def _eval(kind: int, r: torch.Tensor, h: float):
    _lib.require_gpu(r)
    rr = r.detach().contiguous()
    w, g, lap = torch.empty_like(rr), torch.empty_like(rr), torch.empty_like(rr)
    dev = rr.device
    _lib.check(_lib.lib().tsph_kernel_eval(kind, _lib.ptr(rr), rr.numel(), _code(rr), float(h),
                                           _lib.ptr(w), _lib.ptr(g),
                                           _lib.ptr(lap if kind == 1 else None),
                                           _lib.stream_handle(dev)))
    return w, g, lap


def cubic_spline_kernel(r: torch.Tensor, h: float) -> torch.Tensor:
    """W(r, h) with sigma = 1/(pi h^3) (kernel3.py:90-96)."""
    return _eval(1, r, h)[0]


def cubic_spline_gradient(delta: torch.Tensor, h: float, eps: float = 1.0e-12) -> torch.Tensor:
    """grad_i W for delta = x_i - x_j (kernel3.py:99-105)."""
    r = torch.linalg.vector_norm(delta, dim=-1)
    return delta * _eval(1, r, h)[1].unsqueeze(-1)


def cubic_spline_laplacian(delta: torch.Tensor, h: float, eps: float = 1.0e-12) -> torch.Tensor:
    """Radial Laplacian f'' + 2 f'/q scaled by sigma/h^2, -9 sigma/h^2 at r = 0
    (kernel3.py:108-123)."""
    r = torch.linalg.vector_norm(delta, dim=-1)
    return _eval(1, r, h)[2]


def pair_kernels(delta: torch.Tensor, h: float, eps: float = 1.0e-12):
    """(W, grad W, laplacian W) for a batch of separations (kernel3.py:167-181)."""
    r = torch.linalg.vector_norm(delta, dim=-1)
    w, g, lap = _eval(1, r, h)
    return w, delta * g.unsqueeze(-1), lap


def wendland_c2_kernel(r: torch.Tensor, h: float) -> torch.Tensor:
    """3-D Wendland C2, sigma = 21/(16 pi h^3) (kernel3.py:139-148): the 2-D
    kernel of libtsph rescaled by the ratio of the normalisations, 3/(4h)."""
    return _eval(2, r, h)[0] * (3.0 / (4.0 * h))


def wendland_c2_gradient(delta: torch.Tensor, h: float, eps: float = 1.0e-12) -> torch.Tensor:
    """grad_i W of the 3-D Wendland C2 (kernel3.py:151-164)."""
    r = torch.linalg.vector_norm(delta, dim=-1)
    return delta * (_eval(2, r, h)[1] * (3.0 / (4.0 * h))).unsqueeze(-1)


def lattice_kernel_sum(dx: float, h: float, half_width: int | None = None) -> float:
    """C = dx^3 sum W over a simple cubic lattice (kernel3.py:184-198)."""
    return lattice_kernel_sum_3d(dx, h, half_width)


def lattice_bias(dx: float, h: float, half_width: int | None = None) -> float:
    """C - 1 (kernel3.py:201-203)."""
    return lattice_kernel_sum(dx, h, half_width) - 1.0


def calibrated_particle_mass(dx: float, h: float, rho0: float,
                             half_width: int | None = None) -> float:
    """m = rho0 dx^3 / C (kernel3.py:206-215)."""
    return rho0 * dx ** 3 / lattice_kernel_sum(dx, h, half_width)
