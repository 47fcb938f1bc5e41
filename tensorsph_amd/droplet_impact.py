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
                     validate_config, acoustic_courant, adaptive_controls, suggested_dt, stable_dt, next_dt,
                     viscous_dt_limit, gravity_dt_limit, diffusion_dt_limit, acoustic_dt_limit)
from .neighbor_search import PairListCache, cell_linked_pairs, drop_wall_wall_pairs

__all__ = [
    "Config", "SimulationUnstableError", "VISCOUS_LAPLACIAN_FACTOR", "rhs", "tait_pressure",
    "pair_kernels", "cubic_spline_kernel", "cubic_spline_gradient", "wendland_c2_kernel",
    "wendland_c2_gradient", "wall_mass", "drop_wall_wall_pairs", "apply_boundary",
    "boundary_box", "validate_config", "acoustic_courant", "suggested_dt", "stable_dt",
    "next_dt", "lattice_kernel_sum", "lattice_mass_unit", "cell_centres", "domain_top",
    "hydrostatic_density", "PairListCache", "cell_linked_pairs", "run", "evaluate_acceptance",
    "ROW_KEYS",
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


def stepper_options(cfg) -> dict:
    """The Config switches of the rhs that the resident stepper runs too."""
    return {"density_renormalization": bool(cfg.density_renormalization),
            "lattice_kernel_sum": lattice_kernel_sum(cfg.h_factor * cfg.dx, cfg.dx),
            "boundary_consistent_density": bool(cfg.boundary_consistent_density),
            "surface_pressure_closure": bool(cfg.surface_pressure_closure),
            "surface_pressure_threshold": float(cfg.surface_pressure_threshold),
            "surface_pressure_donors": cfg.surface_pressure_donors,
            "surface_pressure_layers": int(cfg.surface_pressure_layers)}


def _rhs_options(p, cfg, wrap_width, dim=2) -> None:
    """The opt-in branches of the 2-D rhs (droplet_impact.py:921-964,
    1031-1112) into the parameter block.  rhs_3d reads only the
    surface-pressure closure of them (dam_break_3d.py:331-335)."""
    if dim == 3:
        if cfg.surface_pressure_closure:
            p.surface_pressure_closure = 1
            p.surface_pressure_threshold = float(cfg.surface_pressure_threshold)
            p.surface_pressure_donors_all = 1 if cfg.surface_pressure_donors == "all" else 0
            p.surface_pressure_layers = int(cfg.surface_pressure_layers)
        return
    p.wrap_width = 0.0 if wrap_width is None else float(wrap_width)
    if cfg.density_renormalization:
        p.density_renormalization = 1
        p.lattice_kernel_sum = lattice_kernel_sum(cfg.h_factor * cfg.dx, cfg.dx)
    p.boundary_consistent_density = 1 if cfg.boundary_consistent_density else 0
    if cfg.surface_pressure_closure:
        p.surface_pressure_closure = 1
        p.surface_pressure_threshold = float(cfg.surface_pressure_threshold)
        p.surface_pressure_donors_all = 1 if cfg.surface_pressure_donors == "all" else 0
        p.surface_pressure_layers = int(cfg.surface_pressure_layers)


def run_wcsph(pos, vel, mass, cfg, pairs, wall, dt, wall_vel, wall_accel, split, dim,
              wall_mass_value, reaction_style, options=None):
    """Shared driver of the 2-D ``rhs`` and the 3-D ``rhs_3d``.  ``options``:
    dict of wrap_width, density_offset, nu_field (2-D rhs); the Config
    branches the rhs of that dimension reads come with it.  With an empty pair list the reference returns
    before every branch, so none is passed then."""
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
    off = nuf = None
    if options is not None and pairs.numel() > 0:
        _rhs_options(prm, cfg, options.get("wrap_width"), dim)
        if options.get("density_offset") is not None:
            off = options["density_offset"].detach().to(sdt).contiguous()
    if options is not None and options.get("nu_field") is not None:
        nuf = options["nu_field"].detach().to(sdt).contiguous()
    rho = torch.empty(n, dtype=sdt, device=dev)
    p = torch.empty(n, dtype=sdt, device=dev)
    acc = torch.empty((n, dim), dtype=sdt, device=dev)
    wall_p = torch.empty(nw, dtype=sdt, device=dev)
    wall_rho = torch.empty(nw, dtype=sdt, device=dev)
    reaction = torch.zeros((nw, dim), dtype=sdt, device=dev) if split else None
    _lib.check(L.tsph_wcsph_rhs_ex(ctx.handle, C.byref(prm), _lib.ptr(all_pos), _lib.ptr(all_vel),
                                   _lib.ptr(all_mass), _lib.ptr(wa), _lib.ptr(off), _lib.ptr(nuf),
                                   _lib.ptr(rho), _lib.ptr(p), _lib.ptr(acc), _lib.ptr(wall_p),
                                   _lib.ptr(wall_rho), _lib.ptr(reaction), stream))
    out_dtype = pos.dtype
    return (rho.to(out_dtype), p.to(out_dtype), acc.to(out_dtype), wall_p.to(out_dtype),
            None if reaction is None else reaction.to(out_dtype))


def rhs(pos, vel, mass, cfg: Config, pairs=None, wall=None, dt=None, wall_vel=None,
        wall_accel=None, state=None, split: bool = False, wrap_width=None, nu_field=None,
        density_offset=None):
    """WCSPH right-hand side on unordered pairs: ``(rho, pressure, acceleration)``.

    With ``split=True`` a fourth value ``{"reaction", "face_pressure"}`` holds
    the force the fluid exerts on each wall particle and the wall pressure.
    The opt-in branches wrap_width, nu_field, density_offset and the Config
    switches density_renormalization, boundary_consistent_density and
    surface_pressure_closure run on the device (droplet_impact.py:921-964,
    1031-1112, 1164-1178)."""
    if pos.shape[1] != 2:
        raise ValueError(f"rhs implements 2-D smoothing only, got {pos.shape[1]} dimensions")
    check_supported(cfg, rhs=True)
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
    if nu_field is not None:
        if not torch.is_tensor(nu_field) or nu_field.shape != (n,):
            raise ValueError(f"nu_field must be a tensor of shape ({n},), got "
                             f"{None if not torch.is_tensor(nu_field) else tuple(nu_field.shape)}")
        if nu_field.device != pos.device or nu_field.dtype != pos.dtype:
            raise ValueError("nu_field must share pos device and dtype")
        if not bool(torch.isfinite(nu_field).all()):
            raise ValueError("nu_field must be finite")
        if bool((nu_field < 0).any()):
            raise ValueError("nu_field must be non-negative")
    if pairs.numel() > 0:
        if wrap_width is not None:
            width = float(wrap_width)
            if not math.isfinite(width) or width <= 0.0:
                raise ValueError("wrap_width must be a positive finite length")
        if density_offset is not None:
            if density_offset.shape != (n,) or density_offset.device != pos.device:
                raise ValueError("density_offset must match the fluid particles")
            if density_offset.dtype != pos.dtype:
                raise ValueError("density_offset must have the fluid dtype")
            if cfg.density_renormalization:
                raise ValueError("density_offset requires unnormalised summation density")
    nw = 0 if wall is None else wall.shape[0]
    rho, p, acc, wall_p, reaction = run_wcsph(
        pos, vel, mass, cfg, pairs, wall, dt, wall_vel, wall_accel, split, 2,
        lattice_mass_unit(cfg), 1,
        {"wrap_width": wrap_width, "density_offset": density_offset, "nu_field": nu_field})
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


# --------------------------------------------------------------------- driver

ROW_KEYS = ["step", "time", "dt", "particle_count", "mass_total", "momentum_x", "momentum_y",
            "rho_min", "rho_max", "rho_mean", "rho_max_relative_error", "pressure_min",
            "pressure_max", "max_speed", "max_acceleration", "mach_number", "courant_number",
            "kinetic_energy", "y_min", "y_max"]


def _meter_row(cfg: Config, m, step: int, n: int) -> dict:
    """One diagnostic row of run() (droplet_impact.py:1711-1733) from a meter
    row of the resident stepper (engine.METER_KEYS: the state of ``step``
    before its integration, as the reference samples it).  An adaptive run
    reports the accumulated time and the size of the step before the row."""
    from .engine import METER_KEYS
    g = {k: float(m[i]) for i, k in enumerate(METER_KEYS)}
    rho_min, rho_max = g["rho_min"], g["rho_max"]
    speed = g["max_speed"]
    time = g["time"] if cfg.adaptive_dt else step * cfg.dt
    last_dt = g["last_dt"] if cfg.adaptive_dt else cfg.dt
    return {
        "step": step, "time": time, "dt": last_dt, "particle_count": n,
        "mass_total": g["mass"], "momentum_x": g["momentum_x"], "momentum_y": g["momentum_y"],
        "rho_min": rho_min, "rho_max": rho_max, "rho_mean": g["rho_sum"] / n,
        "rho_max_relative_error": max(abs(rho_min / cfg.rho0 - 1.0), abs(rho_max / cfg.rho0 - 1.0)),
        "pressure_min": g["pressure_min"], "pressure_max": g["pressure_max"],
        "max_speed": speed, "max_acceleration": g["max_acceleration"],
        "mach_number": speed / cfg.sound_speed, "courant_number": acoustic_courant(cfg.dt, cfg, speed),
        "kinetic_energy": g["kinetic_energy"], "y_min": g["y_min"], "y_max": g["y_max"],
    }


def evaluate_acceptance(rows, cfg: Config, boundary_events: int, extremes: dict | None = None) -> list:
    """droplet_impact.evaluate_acceptance (droplet_impact.py:1456-1512): the
    recorded numerical criteria of a run, from every step's extrema when
    ``extremes`` is given, else from the sampled rows."""
    if extremes is None:
        population = "sampled_rows"
        worst_compression = max(r["rho_max"] / cfg.rho0 - 1.0 for r in rows)
        worst_speed = max(r["max_speed"] for r in rows)
        worst_courant = max(r["courant_number"] for r in rows)
    else:
        population = "all_steps"
        worst_compression = float(extremes["peak_rho_compression"])
        worst_speed = float(extremes["peak_speed"])
        worst_courant = float(extremes["peak_courant"])
    finite = all(math.isfinite(v) for r in rows for v in r.values())

    def criterion(ident, metric, measured, threshold, unit, comparison, scope=population):
        passed = measured >= threshold if comparison == ">=" else measured <= threshold
        return {"id": ident, "metric": metric, "measured": measured, "threshold": threshold,
                "unit": unit, "comparison": comparison, "population": scope,
                "result": "pass" if passed else "fail"}

    out = [criterion("finite_state", "all diagnostics finite", 1.0 if finite else 0.0, 1.0, "bool",
                     ">=", "sampled_rows"),
           criterion("density_compression", "max(rho/rho0 - 1)", worst_compression,
                     cfg.max_density_error, "1", "<="),
           criterion("mach_number", "max|v| / c0", worst_speed / cfg.sound_speed, cfg.max_mach, "1", "<="),
           criterion("acoustic_courant", "max dt(c0+|v|)/h", worst_courant, cfg.max_courant, "1", "<=")]
    escape = criterion("boundary_escape", "particles clamped beyond the box", float(boundary_events), 0.0,
                       "count", "<=", "all_steps")
    if cfg.boundary == "clamp":
        escape["result"] = "not_evaluated"
    out.append(escape)
    return out


def _json_number(v: float):
    return None if not math.isfinite(v) else v


def run(cfg: Config, output=None, plot: bool = True, precision: str | None = None) -> list:
    """droplet_impact.run (droplet_impact.py:1607-1795) on the resident stepper.

    The loop (pairs with the Verlet skin, rhs, every-step extrema, sampled
    rows, symplectic Euler, apply_boundary) runs on the GPU without a host
    round trip per step; rows carry the reference's keys and sampling
    (every ``output_every`` steps and the final one).  With ``output`` the
    diagnostics CSV, the run manifest with the acceptance criteria and one
    particles_XXXXXX.pt snapshot (position, density) per row are written;
    ``plot`` is accepted and ignored (no figures are drawn).  ``precision``
    "fp64" / "fp32" (FP32-compensated) defaults to torch's default dtype, the
    dtype the reference runs in.  With ``adaptive_dt`` the step follows
    next_dt on the device and the run integrates to cfg.steps * cfg.dt (at most
    4 cfg.steps + 1000 steps, as the reference caps it); rows are metered every
    step for the all-step extrema.  ``fail_fast`` raises SimulationUnstableError
    after the artifacts are written, as the reference does.
    """
    import json
    from pathlib import Path

    import numpy as np

    from .cases import make_particles, wall_lattice
    from .engine import WcsphSimulation

    validate_config(cfg)
    check_supported(cfg, rhs=True)
    if precision is None:
        precision = "fp64" if torch.get_default_dtype() == torch.float64 else "fp32"
    pos, mass, n_pool = make_particles(cfg)
    # the box follows the wall height (the harness option ``wall_top``; the
    # reference's box and walls both stop at domain_top)
    x_min, x_max, y_min, y_max = boundary_box(cfg, cfg.wall_top)
    if (pos[:, 0].min() < x_min or pos[:, 0].max() > x_max or pos[:, 1].min() < y_min
            or pos[:, 1].max() > y_max):
        raise ValueError("initial particle arrangement leaves the computational box")
    wall = wall_lattice(cfg, top=cfg.wall_top) if cfg.boundary == "wall" else np.zeros((0, 2))
    vel = np.zeros_like(pos)
    vel[n_pool:] = np.asarray(cfg.initial_velocity, dtype=np.float64)
    h = cfg.h_factor * cfg.dx
    n = pos.shape[0]
    device = "cuda" if cfg.device in ("auto", "cpu") else cfg.device
    sim = WcsphSimulation(pos, vel, mass, wall, h=h, rho0=cfg.rho0, sound_speed=cfg.sound_speed,
                          gamma=cfg.gamma, gravity=cfg.gravity, viscosity=cfg.viscosity, dt=cfg.dt,
                          wall_mass=lattice_mass_unit(cfg), skin=cfg.pair_skin_factor * h,
                          lo=[x_min, y_min], hi=[x_max, y_max], tol=0.25 * cfg.dx,
                          wall_friction=cfg.wall_friction, density_diffusion=cfg.density_diffusion,
                          diffusion_form=cfg.density_diffusion_form,
                          eigenvalue_floor=cfg.density_diffusion_eigenvalue_floor,
                          precision=precision, device=device,
                          adaptive=adaptive_controls(cfg) if cfg.adaptive_dt else None,
                          rhs_options=stepper_options(cfg))
    out = Path(output) if output is not None else None
    if out is not None:
        out.mkdir(parents=True, exist_ok=True)
    snapshots = []
    escaped = 0
    failure = None
    try:
        if cfg.adaptive_dt:
            # the step count is an outcome: advance output_every steps at a time
            # until the final sample (the step starting at the target time)
            max_steps = 4 * cfg.steps + 1000
            step, shot = 0, -1
            while True:
                if out is not None and step != shot:
                    snapshots.append((step, _snapshot(sim, cfg, wall, mass, out, step, time=None)))
                    shot = step
                if sim.stats()["finished"] or step > max_steps:
                    break
                todo = min(cfg.output_every, max_steps + 1 - step)
                sim.step(todo, meter_every=1)
                st = sim.stats()
                step = st["steps"] - 1 if st["finished"] else step + todo
            meters = sim.meters()
            stats = sim.stats()
            escaped = stats["escaped"]
            if not stats["finished"]:
                failure = (f"the adaptive step did not reach {cfg.steps * cfg.dt:g} s within "
                           f"{max_steps} steps")
            final_step = int(meters[-1, 0]) if stats["finished"] else -1
            row_steps = [int(r[0]) for r in meters
                         if int(r[0]) % cfg.output_every == 0 or int(r[0]) == final_step]
        else:
            row_steps = list(range(0, cfg.steps, cfg.output_every)) + [cfg.steps]
            for k, step in enumerate(row_steps):
                if out is not None:
                    snapshots.append((step, _snapshot(sim, cfg, wall, mass, out, step)))
                if step == cfg.steps:
                    escaped = sim.stats()["escaped"]
                    sim.step(1, meter_every=1)         # the final row: rhs of the last state
                    break
                nxt = row_steps[k + 1]
                sim.step(nxt - step, meter_every=cfg.output_every)
            meters = sim.meters()
            stats = sim.stats()
    finally:
        sim.close()
    by_step = {int(r[0]): r for r in meters}
    rows = [_meter_row(cfg, by_step[s], s, n) for s in row_steps if s in by_step]
    peak_speed = float(stats["peak_speed"])
    if cfg.adaptive_dt:
        from .engine import METER_KEYS
        col = {k: i for i, k in enumerate(METER_KEYS)}
        peak_courant = max(acoustic_courant(float(r[col["dt"]]), cfg, float(r[col["max_speed"]]))
                           for r in meters)
        last = meters[-1]
        simulated_time = float(last[col["time"]] + last[col["step_dt"]])
        dt_min, dt_max = float(meters[:, col["dt"]].min()), float(meters[:, col["dt"]].max())
    else:
        peak_courant = acoustic_courant(cfg.dt, cfg, peak_speed)
        simulated_time, dt_min, dt_max = cfg.steps * cfg.dt, cfg.dt, cfg.dt
    extremes = {"peak_rho_compression": float(stats["peak_rho"]) / cfg.rho0 - 1.0,
                "peak_speed": peak_speed, "peak_courant": peak_courant,
                "simulated_time": simulated_time, "dt_min": dt_min, "dt_max": dt_max}
    criteria = evaluate_acceptance(rows, cfg, escaped, extremes=extremes)
    if failure is None and not all(math.isfinite(v) for r in rows for v in r.values()):
        failure = "non-finite state"
    if out is not None:
        import csv
        import dataclasses
        with (out / "diagnostics.csv").open("w", newline="") as f:
            writer = csv.DictWriter(f, fieldnames=ROW_KEYS)
            writer.writeheader()
            writer.writerows(rows)
        numerical = "fail" if any(c["result"] == "fail" for c in criteria) else "pass"
        manifest = {
            "case": "droplet_impact_2d_wcsph",
            "solver": {"name": "tensorsph_amd", "device": str(torch.device(device)),
                       "precision": precision},
            "initial_particle_count": int(n), "wall_particle_count": int(wall.shape[0]),
            "config": {k: (list(v) if isinstance(v, tuple) else v)
                       for k, v in dataclasses.asdict(cfg).items()},
            "derived": {"smoothing_length": h, "suggested_dt": suggested_dt(cfg),
                        "viscous_dt_limit": _json_number(viscous_dt_limit(cfg)),
                        "gravity_dt_limit": _json_number(gravity_dt_limit(cfg)),
                        "diffusion_dt_limit": _json_number(diffusion_dt_limit(cfg)),
                        "acoustic_courant": acoustic_courant(cfg.dt, cfg),
                        "simulated_time": extremes["simulated_time"], "adaptive_dt": bool(cfg.adaptive_dt),
                        "dt_min": dt_min, "dt_max": dt_max},
            "statuses": {"software_execution": "failed" if failure else "success",
                         "numerical_verification": numerical, "physical_validation": "not_evaluated"},
            "acceptance": {"criteria": criteria},
            "performance": {"pair_skin": cfg.pair_skin_factor * h, "pair_rebuilds": int(stats["rebuilds"])},
            "sampling": {"diagnostic_rows": len(rows), "output_every": cfg.output_every,
                         "extrema_population": "all_steps"},
            "extremes": extremes, "boundary_events": int(escaped),
            "snapshots": [f"particles_{s:06d}.pt" for s, _ in snapshots],
        }
        (out / "run_manifest.json").write_text(json.dumps(manifest, indent=2))
    if cfg.fail_fast:
        failed = [c for c in criteria if c["result"] == "fail"]
        if failure or failed:
            detail = failure or ", ".join(f"{c['id']}={c['measured']:.4g} (limit {c['threshold']:g})"
                                          for c in failed)
            raise SimulationUnstableError(f"numerical acceptance failed: {detail}"
                                          + (f"; artifacts in {out}" if out is not None else ""))
    return rows


def _snapshot(sim, cfg: Config, wall, mass, out, step: int, time: float | None = None):
    """particles_XXXXXX.pt of the current state: positions and the rhs density
    of that state (droplet_impact.py:1734-1739).  ``time`` defaults to
    step * dt (fixed step); an adaptive run passes None for the accumulated
    time of the state, read from the stepper."""
    if time is None and cfg.adaptive_dt:
        time = _state_time(sim)
    elif time is None:
        time = step * cfg.dt
    st = sim.read()
    pos = st["position"]
    dt = pos.dtype if sim.code == _lib.F64 else torch.float32
    p = pos.to(dt)
    w = torch.as_tensor(wall, dtype=dt, device=pos.device) if len(wall) else None
    rho, _, _ = rhs(p, st["velocity"].to(dt), torch.as_tensor(mass, dtype=dt, device=pos.device), cfg,
                    wall=w, dt=cfg.dt)
    name = out / f"particles_{step:06d}.pt"
    torch.save({"step": int(step), "time": float(time), "position": pos.detach().cpu(),
                "density": rho.detach().cpu(), "boundary": str(cfg.boundary)}, name)
    return name


def _state_time(sim) -> float:
    """Simulated time of the stepper's current state: the last metered step's
    start time plus its size (0 before the first step)."""
    from .engine import METER_KEYS
    m = sim.meters()
    if m.shape[0] == 0:
        return 0.0
    col = {k: i for i, k in enumerate(METER_KEYS)}
    return float(m[-1, col["time"]] + m[-1, col["step_dt"]])
