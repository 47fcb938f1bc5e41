"""2-D WCSPH operators behind the reference's ``droplet_impact`` API.

``rhs`` (droplet_impact.py:810-1228), ``tait_pressure`` (:415-427),
``pair_kernels`` / ``cubic_spline_kernel`` / ``cubic_spline_gradient``
(:276-371), ``wall_mass`` (:747-754) and ``apply_boundary`` (:1293-1312),
computed by libtsph on the GPU.  Same arguments, same return values and the
same exceptions as the reference; tensors must live on a ROCm device.

Precision follows the positions' dtype: float64 runs the plain FP64
kernels; float32 (the reference's default 2-D dtype) runs the FP32
compensated kernels (pair deltas formed in FP64, kernel arithmetic in FP32,
Kahan-compensated sums) and returns float32.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from . import _lib
from .cases import (cell_centres, domain_top, hydrostatic_density, lattice_kernel_sum,
                    lattice_mass_unit)
from .config import (Config, SimulationUnstableError, VISCOUS_LAPLACIAN_FACTOR, check_supported,
                     validate_config, acoustic_courant, suggested_dt, stable_dt, next_dt,
                     viscous_dt_limit, gravity_dt_limit, diffusion_dt_limit, acoustic_dt_limit)
from .neighbor_search import PairListCache, cell_linked_pairs, drop_wall_wall_pairs

__all__ = [
    "Config", "SimulationUnstableError", "VISCOUS_LAPLACIAN_FACTOR", "rhs", "tait_pressure",
    "pair_kernels", "cubic_spline_kernel", "cubic_spline_gradient", "wendland_c2_kernel",
    "wendland_c2_gradient", "wall_mass", "drop_wall_wall_pairs", "apply_boundary",
    "boundary_box", "validate_config", "acoustic_courant", "suggested_dt", "stable_dt",
    "next_dt", "lattice_kernel_sum", "lattice_mass_unit", "cell_centres", "domain_top",
    "hydrostatic_density", "PairListCache", "cell_linked_pairs",
]


def _precision(t: torch.Tensor) -> int:
    if t.dtype == torch.float64:
        return _lib.F64
    if t.dtype == torch.float32:
        return _lib.F32
    raise TypeError(f"positions must be float32 or float64, got {t.dtype}")


def _scalar_dtype(code: int) -> torch.dtype:
    return torch.float64 if code == _lib.F64 else torch.float32


# --------------------------------------------------------------------- kernels

def _kernel_eval(kind: int, r: torch.Tensor, h: float, with_grad: bool, with_lap: bool = False):
    dev = _lib.require_gpu(r)
    code = _precision(r)
    rr = r.detach().contiguous()
    w = torch.empty_like(rr)
    g = torch.empty_like(rr) if with_grad else None
    lap = torch.empty_like(rr) if with_lap else None
    _lib.check(_lib.lib().tsph_kernel_eval(kind, _lib.ptr(rr), rr.numel(), code, float(h),
                                           _lib.ptr(w), _lib.ptr(g), _lib.ptr(lap),
                                           _lib.stream_handle(dev)))
    return w, g, lap


def cubic_spline_kernel(r: torch.Tensor, h: float) -> torch.Tensor:
    """2-D cubic spline W(r, h), sigma = 10/(7 pi h^2)."""
    return _kernel_eval(0, r, h, False)[0]


def _norm(delta: torch.Tensor) -> torch.Tensor:
    return torch.linalg.vector_norm(delta, dim=-1)


def cubic_spline_gradient(delta: torch.Tensor, h: float, eps: float = 1.0e-12) -> torch.Tensor:
    """grad_i W for delta = x_i - x_j."""
    _, g, _ = _kernel_eval(0, _norm(delta), h, True)
    return delta * g.unsqueeze(-1)


def wendland_c2_kernel(r: torch.Tensor, h: float) -> torch.Tensor:
    return _kernel_eval(2, r, h, False)[0]


def wendland_c2_gradient(delta: torch.Tensor, h: float, eps: float = 1.0e-12) -> torch.Tensor:
    _, g, _ = _kernel_eval(2, _norm(delta), h, True)
    return delta * g.unsqueeze(-1)


def pair_kernels(delta: torch.Tensor, h: float, eps: float = 1.0e-12):
    """``(W, grad W)`` for a batch of pair separations (droplet_impact.py:349-371)."""
    w, g, _ = _kernel_eval(0, _norm(delta), h, True)
    return w, delta * g.unsqueeze(-1)


def tait_pressure(rho: torch.Tensor, cfg: Config) -> torch.Tensor:
    """Tait EOS with the ratio floored at 0.01 and p clamped at 0."""
    dev = _lib.require_gpu(rho)
    code = _precision(rho)
    r = rho.detach().contiguous()
    p = torch.empty_like(r)
    _lib.check(_lib.lib().tsph_eos_tait(_lib.ptr(r), r.numel(), code, cfg.rho0, cfg.sound_speed,
                                        cfg.gamma, 0.01, _lib.ptr(p), _lib.stream_handle(dev)))
    return p


def wall_mass(cfg: Config, count: int, device) -> torch.Tensor:
    return torch.full((count,), lattice_mass_unit(cfg), device=device)


# --------------------------------------------------------------------- rhs

def _wcsph_params(cfg, n, nw, dim, code, step_dt, reaction_style, has_wall_accel, diffusion):
    p = _lib.WcsphParams()
    p.dim = dim
    p.precision = code
    p.n_fluid = n
    p.n_wall = nw
    p.h = cfg.h_factor * cfg.dx
    p.rho0 = cfg.rho0
    p.sound_speed = cfg.sound_speed
    p.gamma = cfg.gamma
    p.gravity = cfg.gravity
    p.viscosity = cfg.viscosity
    p.wall_friction = cfg.wall_friction
    p.dt = step_dt
    if diffusion and cfg.density_diffusion > 0.0:
        p.diffusion_scale = 2.0 * cfg.density_diffusion * cfg.sound_speed * cfg.h_factor * cfg.dx
        p.diffusion_form = 1 if cfg.density_diffusion_form == "antuono" else 0
    p.eigenvalue_floor = cfg.density_diffusion_eigenvalue_floor
    p.reaction_style = reaction_style
    p.has_wall_accel = 1 if has_wall_accel else 0
    return p


def run_wcsph(pos, vel, mass, cfg, pairs, wall, dt, wall_vel, wall_accel, split, dim,
              wall_mass_value, reaction_style):
    """Shared driver of the 2-D ``rhs`` and the 3-D ``rhs_3d``."""
    dev = _lib.require_gpu(pos, vel, mass, wall, pairs)
    code = _precision(pos)
    sdt = _scalar_dtype(code)
    n = pos.shape[0]
    nw = 0 if wall is None else wall.shape[0]
    all_pos = pos if wall is None else torch.cat((pos, wall), dim=0)
    all_pos = all_pos.detach().to(torch.float64).contiguous()
    if nw:
        wv = torch.zeros((nw, dim), dtype=torch.float64, device=dev) if wall_vel is None \
            else wall_vel.detach().to(torch.float64)
        all_vel = torch.cat((vel.detach().to(torch.float64), wv), dim=0).contiguous()
        all_mass = torch.cat((mass.detach().to(sdt),
                              torch.full((nw,), wall_mass_value, dtype=sdt, device=dev))).contiguous()
    else:
        all_vel = vel.detach().to(torch.float64).contiguous()
        all_mass = mass.detach().to(sdt).contiguous()
    pr = pairs.detach().contiguous()
    ctx = _lib.default_context(dev)
    stream = _lib.stream_handle(dev)
    L = _lib.lib()
    _lib.check(L.tsph_nlist_from_pairs(ctx.handle, _lib.ptr(pr), pr.shape[0], n, n + nw, 0, stream))
    wa = None
    if wall_accel is not None:
        if wall_accel.shape != (nw, dim):
            raise ValueError("wall_accel must be (wall_count, pos.shape[1])")
        if wall_accel.device != pos.device or wall_accel.dtype != pos.dtype:
            raise ValueError("wall_accel must share pos device and dtype")
        wa = wall_accel.detach().to(torch.float64).contiguous()
    step_dt = cfg.dt if dt is None else float(dt)
    prm = _wcsph_params(cfg, n, nw, dim, code, step_dt, reaction_style if split else 0,
                        wa is not None, True)
    rho = torch.empty(n, dtype=sdt, device=dev)
    p = torch.empty(n, dtype=sdt, device=dev)
    acc = torch.empty((n, dim), dtype=sdt, device=dev)
    wall_p = torch.empty(nw, dtype=sdt, device=dev)
    wall_rho = torch.empty(nw, dtype=sdt, device=dev)
    reaction = torch.zeros((nw, dim), dtype=sdt, device=dev) if split else None
    _lib.check(L.tsph_wcsph_rhs(ctx.handle, C.byref(prm), _lib.ptr(all_pos), _lib.ptr(all_vel),
                                _lib.ptr(all_mass), _lib.ptr(wa), _lib.ptr(rho), _lib.ptr(p),
                                _lib.ptr(acc), _lib.ptr(wall_p), _lib.ptr(wall_rho),
                                _lib.ptr(reaction), stream))
    out_dtype = pos.dtype
    return (rho.to(out_dtype), p.to(out_dtype), acc.to(out_dtype), wall_p.to(out_dtype),
            None if reaction is None else reaction.to(out_dtype))


def rhs(pos, vel, mass, cfg: Config, pairs=None, wall=None, dt=None, wall_vel=None,
        wall_accel=None, state=None, split: bool = False, wrap_width=None, nu_field=None,
        density_offset=None):
    """WCSPH right-hand side on unordered pairs: ``(rho, pressure, acceleration)``.

    With ``split=True`` a fourth value ``{"reaction", "face_pressure"}`` holds
    the force the fluid exerts on each wall particle and the wall pressure."""
    if pos.shape[1] != 2:
        raise ValueError(f"rhs implements 2-D smoothing only, got {pos.shape[1]} dimensions")
    for name, val in (("wrap_width", wrap_width), ("nu_field", nu_field),
                      ("density_offset", density_offset)):
        if val is not None:
            raise NotImplementedError(f"rhs({name}=...) is an opt-in branch outside the GPU "
                                      "particle pipeline")
    check_supported(cfg)
    if cfg.density_diffusion > 0.0 and cfg.density_diffusion_form not in ("molteni", "antuono"):
        raise ValueError(f"unknown density_diffusion_form {cfg.density_diffusion_form!r}")
    h = cfg.h_factor * cfg.dx
    n = pos.shape[0]
    if wall is not None:
        if wall.ndim != 2 or wall.shape[1] != pos.shape[1]:
            raise ValueError("wall must be (W, D) with the same D as pos")
        if wall.device != pos.device or wall.dtype != pos.dtype:
            raise ValueError("wall must share pos device and dtype")
    if wall_vel is not None and wall is not None:
        if wall_vel.shape != (wall.shape[0], pos.shape[1]):
            raise ValueError("wall_vel must be (wall_count, pos.shape[1])")
        if wall_vel.device != pos.device or wall_vel.dtype != pos.dtype:
            raise ValueError("wall_vel must share pos device and dtype")
    _lib.require_gpu(pos, vel, mass, wall)
    if pairs is None:
        all_pos = pos if wall is None else torch.cat((pos, wall), dim=0)
        pairs = cell_linked_pairs(all_pos, 2.0 * h)
        pairs = drop_wall_wall_pairs(pairs, n)
    if pairs.ndim != 2 or pairs.shape[1] != 2 or pairs.dtype != torch.long:
        raise ValueError("pairs must be a long tensor with shape (pair_count, 2)")
    if pairs.device != pos.device:
        raise ValueError("pairs and pos must be on the same device")
    nw = 0 if wall is None else wall.shape[0]
    rho, p, acc, wall_p, reaction = run_wcsph(
        pos, vel, mass, cfg, pairs, wall, dt, wall_vel, wall_accel, split, 2,
        lattice_mass_unit(cfg), 1)
    if pairs.numel() == 0:
        return rho, p, acc
    if split:
        return rho, p, acc, {"reaction": reaction, "face_pressure": wall_p}
    return rho, p, acc


# --------------------------------------------------------------------- boundary

def boundary_box(cfg: Config, top: float | None = None):
    height = domain_top(cfg) if top is None else float(top)
    if cfg.boundary == "wall":
        return 0.0, cfg.pool_width, 0.0, height
    margin = 0.5 * cfg.dx
    return margin, cfg.pool_width - margin, margin, height


def apply_boundary(pos, vel, cfg: Config, top: float | None = None) -> int:
    """Safety clamp (droplet_impact.py:1293-1312), in place; returns the crossings."""
    from .engine import boundary_clamp
    x_min, x_max, y_min, y_max = boundary_box(cfg, top)
    return boundary_clamp(pos, vel, [x_min, y_min], [x_max, y_max], 0.25 * cfg.dx)
