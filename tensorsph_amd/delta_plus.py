"""δ⁺-SPH step behind the reference's ``delta_plus`` API (delta_plus.py:19-229).

``delta_plus_step`` keeps the reference's signature, validation, exceptions and
``DeltaPlusStep`` result; the step itself is the four FP64 sweeps of libtsph
(tsph_deltaplus_step) over neighbour lists built on the GPU from the caller's
pair list.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import torch

from . import _lib
from .neighbor_search import cell_linked_pairs

__all__ = ["DeltaPlusStep", "delta_plus_step"]


@dataclass
class DeltaPlusStep:
    position: torch.Tensor
    velocity: torch.Tensor
    density: torch.Tensor
    acceleration: torch.Tensor
    wall_reaction: torch.Tensor
    shift: torch.Tensor


def delta_plus_step(pos, velocity, density, mass, *, dx, sound_speed, reference_speed, dt,
                    rho0=1000., gravity=9.81, viscosity=1.e-6, diffusion=.1, alpha=.02,
                    wall=None, wall_velocity=None, wall_normal=None, pairs=None):
    """Advance an equal-spacing 2-D liquid by one δ⁺-SPH step (delta_plus.py:105-229).

    Smoothing length 2 dx; positions advance with the new velocity plus one
    positional shift; wall stations of volume dx^2 with slip ghosts.  Pairs
    follow the package convention (i < j over cat(pos, wall)); cached lists
    may include skin pairs.  The step runs in FP64: float32 state is widened
    for it (the pair search, when pairs is None, runs in float32 as the
    reference's does) and the result is returned as float32."""
    if (pos.ndim != 2 or pos.shape[1] != 2 or velocity.shape != pos.shape
            or density.shape != pos.shape[:1] or mass.shape != density.shape
            or not len(pos) or pos.dtype not in (torch.float32, torch.float64)):
        raise ValueError("expected a nonempty 2-D fluid and one density and mass per particle")
    if not all(math.isfinite(x) for x in (dx, sound_speed, reference_speed, dt,
                                          rho0, gravity, viscosity, diffusion, alpha)):
        raise ValueError("solver parameters must be finite")
    if min(dx, sound_speed, dt, rho0) <= 0 or min(reference_speed, viscosity, diffusion, alpha) < 0:
        raise ValueError("invalid length, speed, density, step or transport coefficient")
    h = 2 * dx
    radial_viscosity = 8 * viscosity + alpha * h * sound_speed
    if (dt * sound_speed > .3 * h or dt * radial_viscosity / 8 > .125 * h * h
            or dt * diffusion * sound_speed > .125 * h):
        raise ValueError("time step exceeds an acoustic, viscous or diffusive limit")
    if wall is None:
        wall = pos.new_empty((0, 2))
    if wall_velocity is None:
        wall_velocity = torch.zeros_like(wall)
    if wall.ndim != 2 or wall.shape[1] != 2 or wall_velocity.shape != wall.shape:
        raise ValueError("expected matching (W, 2) wall positions and velocities")
    if wall_normal is None and len(wall) == 0:
        wall_normal = wall
    if wall_normal is None or wall_normal.shape != wall.shape:
        raise ValueError("provide a wall_normal per ghost station for free slip")
    if any(x.dtype != pos.dtype or x.device != pos.device
           for x in (velocity, density, mass, wall, wall_velocity, wall_normal)):
        raise ValueError("states must share a floating dtype and device")
    dev = _lib.require_gpu(pos)
    if not bool(((density > 0) & (mass > 0) & density.isfinite() & mass.isfinite()).all()):
        raise ValueError("density and mass must be positive and finite")
    norms = wall_normal.norm(dim=1)
    if not bool((norms.isfinite() & (norms > 0)).all()):
        raise ValueError("wall normals must be finite and nonzero")
    wall_normal = wall_normal / norms[:, None]
    if not all(bool(x.isfinite().all()) for x in (pos, velocity, wall, wall_velocity)):
        raise ValueError("particle positions and velocities must be finite")
    n, nw = pos.shape[0], wall.shape[0]
    if pairs is None:
        pairs = cell_linked_pairs(torch.cat((pos, wall)), 2 * h)
    if pairs.ndim != 2 or pairs.shape[1] != 2 or pairs.dtype != torch.long:
        raise ValueError("expected a (P, 2) long pair list")
    if pairs.device != pos.device:
        raise ValueError("pair list must share the device of the state")
    if pairs.numel():
        i, j = pairs.unbind(1)
        if bool(((i < 0) | (i >= j) | (j >= n + nw)).any()):
            raise ValueError("each pair must have 0 <= i < j < number of particles")
    out_dtype = pos.dtype
    if out_dtype != torch.float64:
        pos, velocity, density, mass, wall, wall_velocity, wall_normal = (
            x.to(torch.float64) for x in (pos, velocity, density, mass, wall, wall_velocity,
                                          wall_normal))
    ctx = _lib.default_context(dev)
    stream = _lib.stream_handle(dev)
    L = _lib.lib()
    pr = pairs.contiguous()
    _lib.check(L.tsph_nlist_from_pairs(ctx.handle, _lib.ptr(pr), pr.shape[0], n, n + nw, 0, stream))
    prm = _lib.DeltaPlusParams()
    prm.precision, prm.n, prm.n_wall = _lib.F64, n, nw
    prm.dx, prm.sound_speed, prm.reference_speed, prm.dt = dx, sound_speed, reference_speed, dt
    prm.rho0, prm.gravity, prm.viscosity, prm.diffusion, prm.alpha = rho0, gravity, viscosity, \
        diffusion, alpha
    c = [t.contiguous() for t in (pos, velocity, density, mass, wall, wall_velocity, wall_normal)]
    out = DeltaPlusStep(torch.empty_like(pos), torch.empty_like(pos), torch.empty_like(density),
                        torch.empty_like(pos), torch.zeros_like(wall), torch.empty_like(pos))
    _lib.check(L.tsph_deltaplus_step(ctx.handle, C.byref(prm), *[_lib.ptr(t) for t in c],
                                     _lib.ptr(out.position), _lib.ptr(out.velocity),
                                     _lib.ptr(out.density), _lib.ptr(out.acceleration),
                                     _lib.ptr(out.wall_reaction if nw else None),
                                     _lib.ptr(out.shift), stream))
    if out_dtype != torch.float64:
        out = DeltaPlusStep(*(x.to(out_dtype) for x in (out.position, out.velocity, out.density,
                                                          out.acceleration, out.wall_reaction,
                                                          out.shift)))
    return out
