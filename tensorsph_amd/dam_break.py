"""2-D dam break behind the reference's ``dam_break`` API (dam_break.py:45-486).

``run_case`` integrates the water column on the resident stepper (Verlet
pair cache, rhs, symplectic Euler, apply_boundary: the loop of
dam_break.py:374-465) and returns the same ``DamBreakRun`` record: the
sampled front position and energies, every-step extrema, the wall-impact
time and the compression before it.  Lind shifting (``shift_every``) and the
transport-velocity correction (``tvc_every``) run on the device after the
steps they follow (tsph_sim_shift, tsph_sim_tvc); the adaptive step runs on
the device.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from .cases import make_column, wall_lattice, lattice_mass_unit
from .config import DamBreakConfig, adaptive_controls, check_supported

__all__ = ["DamBreakConfig", "DamBreakRun", "validate_case", "front_position", "ritter_front",
           "ritter_height_profile", "ritter_front_over_a", "run_case"]


@dataclass
class DamBreakRun:
    """The record of dam_break.DamBreakRun (dam_break.py:234-265)."""
    times: list
    front: list
    kinetic_energy: list
    potential_energy: list
    mass: float
    escaped: int
    rebuilds: int
    max_speed: float
    max_density_compression: float
    compression_before_wall_impact: float
    wall_impact_time: float | None
    initial_mass: float
    final_mass: float
    final_positions: torch.Tensor
    steps_taken: int = 0
    simulated_time: float = 0.0
    dt_min: float = 0.0
    dt_max: float = 0.0
    max_acoustic_courant: float = 0.0
    max_tvc_shift: float = 0.0
    max_tvc_clipped: float = 0.0
    snapshots: list = field(default_factory=list)

    @property
    def mass_drift(self) -> float:
        """Final minus initial total mass (dam_break.py:271-279): particle masses
        are constants, so this is zero, measured rather than assumed."""
        return self.final_mass - self.initial_mass

    def dimensionless(self, cfg: DamBreakConfig):
        """(tau, Z/a) samples with tau = t sqrt(2 g / a), a the column width
        (dam_break.py:281-284)."""
        a = cfg.column_width
        rate = math.sqrt(2.0 * cfg.gravity / a)
        return [t * rate for t in self.times], [z / a for z in self.front]


def validate_case(cfg: DamBreakConfig) -> None:
    """dam_break.validate_case (dam_break.py:123-140)."""
    if cfg.boundary != "wall":
        raise ValueError("the dam-break case integrates static wall particles only; "
                         f"boundary={cfg.boundary!r} would be silently ignored")
    if cfg.dx <= 0 or cfg.h_factor <= 0:
        raise ValueError("dx and h_factor must be positive")
    if cfg.column_width <= 0 or cfg.column_height <= 0:
        raise ValueError("column dimensions must be positive")
    if cfg.column_width > cfg.pool_width or cfg.column_height > cfg.tank_height:
        raise ValueError("the column must fit inside the tank")
    if cfg.sound_speed <= 0 or cfg.rho0 <= 0 or cfg.gamma <= 0:
        raise ValueError("rho0, sound_speed, and gamma must be positive")
    courant = cfg.dt * cfg.sound_speed / (cfg.h_factor * cfg.dx)
    if courant > cfg.max_courant * (1.0 + 1.0e-9):
        raise ValueError(f"dt={cfg.dt:g} gives acoustic Courant {courant:.3f} > max_courant "
                         f"{cfg.max_courant:g}; use dt <= {cfg.dt_from_courant():.3e} s")


def front_position(pos, cfg: DamBreakConfig) -> float:
    """Leading-edge x of the water (dam_break.py:172-176)."""
    if pos.numel() == 0:
        return 0.0
    return float(pos[:, 0].max())


def ritter_front(time: float, cfg: DamBreakConfig) -> float:
    """Ritter's dry-bed front distance from the column face, 2 sqrt(g H) t, with
    H the column height (dam_break.py:179-196).  Compare it with
    front_position - column_width."""
    return 2.0 * math.sqrt(cfg.gravity * cfg.column_height) * time


def ritter_height_profile(x: float, time: float, cfg: DamBreakConfig) -> float | None:
    """Ritter's free surface h(x, t), x measured from the column face
    (dam_break.py:199-219): H behind the rarefaction (x < -c t), the fan
    (2c - x/t)^2 / (9 g) for -c t <= x <= 2 c t, dry beyond; c = sqrt(g H).
    None before release (t <= 0)."""
    if time <= 0.0:
        return None
    c = math.sqrt(cfg.gravity * cfg.column_height)
    if x < -c * time:
        return cfg.column_height
    if x > 2.0 * c * time:
        return 0.0
    return (2.0 * c - x / time) ** 2 / (9.0 * cfg.gravity)


def ritter_front_over_a(tau: float) -> float:
    """The Ritter front in Z/a against tau = t sqrt(2 g / a) for H = 2a: 2 tau
    (dam_break.py:222-230)."""
    return 2.0 * tau


def run_case(cfg: DamBreakConfig, pair_skin_factor: float | None = None, snapshot_times=None,
             precision: str | None = None, device=None) -> DamBreakRun:
    """dam_break.run_case (dam_break.py:316-486) on the resident stepper.

    Every step is metered on the device (its pre-integration state, as the
    reference samples it), so the all-step extrema and the first step whose
    front reaches pool_width - 3 dx come from the same rows; samples every
    ``output_every`` steps and at the end, at the accumulated simulated time.
    ``snapshot_times`` captures the fluid positions and velocities at the first
    sample reaching each time.  With ``adaptive_dt`` the step follows next_dt
    on the device up to cfg.steps * cfg.dt (at most 4 cfg.steps + 1000 steps).
    ``precision`` "fp64" / "fp32" (FP32-compensated) defaults to torch's
    default dtype."""
    from .engine import METER_KEYS, WcsphSimulation
    from .droplet_impact import _state_time, stepper_options

    validate_case(cfg)
    check_supported(cfg, rhs=True, shifting=True)
    se = int(cfg.shift_every or 0)
    te = int(cfg.tvc_every or 0)
    tvc_kw = None
    if te:
        name = cfg.tvc_limiter
        if name is None or str(name).lower() in ("none", "no_limiter", "nolimiter"):
            slope = 0.0
        elif str(name).lower() in ("truncated_linear", "truncatedlinear"):
            slope = float(cfg.tvc_limiter_slope)
            if not math.isfinite(slope) or slope <= 0.0:
                raise ValueError("TruncatedLinear slope must be finite and positive")
        else:
            raise ValueError(f"unknown tvc_limiter {name!r}")
        if cfg.tvc_scope == "bulk_g" and (not math.isfinite(cfg.tvc_bulk_g) or cfg.tvc_bulk_g < 0.0):
            raise ValueError("scope_tol must be finite and non-negative")
        tvc_kw = dict(alpha=0.05 if cfg.tvc_alpha is None else float(cfg.tvc_alpha), limiter_slope=slope,
                      max_shift=None if cfg.tvc_max_shift is None else float(cfg.tvc_max_shift) * cfg.h_factor * cfg.dx,
                      wall_volume=cfg.dx ** 2, scope=cfg.tvc_scope, bulk_c=float(cfg.tvc_bulk_c),
                      bulk_g=float(cfg.tvc_bulk_g))
    tvc_peak = tvc_clip_peak = 0.0
    if precision is None:
        precision = "fp64" if torch.get_default_dtype() == torch.float64 else "fp32"
    dev = torch.device(device) if device is not None else torch.device("cuda")
    top = cfg.tank_top()
    pos, mass = make_column(cfg)
    wall = wall_lattice(cfg, top=top)
    h = cfg.h_factor * cfg.dx
    skin = cfg.pair_skin_factor if pair_skin_factor is None else pair_skin_factor
    initial_mass = float(mass.sum())
    adaptive = bool(cfg.adaptive_dt)
    sim = WcsphSimulation(pos, np.zeros_like(pos), mass, wall, h=h, rho0=cfg.rho0,
                          sound_speed=cfg.sound_speed, gamma=cfg.gamma, gravity=cfg.gravity,
                          viscosity=cfg.viscosity, dt=cfg.dt, wall_mass=lattice_mass_unit(cfg),
                          skin=skin * h, lo=[0.0, 0.0], hi=[cfg.pool_width, top], tol=0.25 * cfg.dx,
                          wall_friction=cfg.wall_friction, density_diffusion=cfg.density_diffusion,
                          diffusion_form=cfg.density_diffusion_form,
                          eigenvalue_floor=cfg.density_diffusion_eigenvalue_floor,
                          precision=precision, device=dev,
                          adaptive=adaptive_controls(cfg) if adaptive else None,
                          rhs_options=stepper_options(cfg))
    targets = sorted(float(t) for t in (snapshot_times or []))
    snapshots = []
    max_steps = 4 * cfg.steps + 1000 if adaptive else cfg.steps
    ow = cfg.output_every
    try:
        step = 0                                   # the step whose start state the stepper holds
        while True:
            final = sim.stats()["finished"] if adaptive else step == cfg.steps
            # a snapshot is taken at the first sample reaching its time (dam_break.py:410-414)
            if targets and (step % ow == 0 or final):
                t_state = _state_time(sim)
                while targets and t_state >= targets[0] - 1.0e-12:
                    st = sim.read()
                    snapshots.append((t_state, st["position"].clone(), st["velocity"].clone()))
                    targets.pop(0)
            if final or step > max_steps:
                break
            nxt = min((step // ow + 1) * ow, max_steps + 1)
            for every in (se, te):
                if every:
                    # stop after each step s > 0 with s % every == 0 (not the last): the
                    # corrections move the fluid before step s + 1 (dam_break.py:424-460)
                    s = ((max(step, 1) + every - 1) // every) * every
                    if s < (max_steps if adaptive else cfg.steps):
                        nxt = min(nxt, s + 1)
            sim.step(nxt - step, meter_every=1)
            last = nxt - 1
            moved = (not sim.stats()["finished"]) if adaptive else last < cfg.steps
            if (se or te) and last > 0 and moved:
                # the step's own dt: 0 asks the stepper for its last step's size
                step_dt = 0.0 if adaptive else cfg.dt
                if se and last % se == 0:
                    sim.shift(step_dt, v_ref=cfg.shift_v_ref, v_floor=0.05 * cfg.sound_speed,
                              strength=cfg.shift_strength, limit=cfg.shift_limit)
                if te and last % te == 0:
                    info = sim.tvc(**tvc_kw)
                    tvc_peak = max(tvc_peak, info["max_shift"])
                    tvc_clip_peak = max(tvc_clip_peak, info["clipped"])
            if adaptive and sim.stats()["finished"]:
                step = sim.stats()["steps"] - 1
            else:
                step = nxt
        escaped = sim.stats()["escaped"]
        final_pos = sim.read()["position"]
        if not adaptive:
            sim.step(1, meter_every=1)             # the final sample: rhs of the last state
        rows = sim.meters()
        stats = sim.stats()
    finally:
        sim.close()
    if adaptive and not stats["finished"]:
        raise FloatingPointError(f"the adaptive step did not reach {cfg.steps * cfg.dt:g} s within "
                                 f"{max_steps} steps")
    col = {k: i for i, k in enumerate(METER_KEYS)}
    rows = rows[np.argsort(rows[:, 0], kind="stable")]
    steps_taken = int(rows[-1, 0]) if adaptive else cfg.steps
    rows = rows[: steps_taken + 1]
    sample = [k for k in range(steps_taken + 1) if k % ow == 0 or k == steps_taken]
    time = rows[:, col["time"]]
    impact_limit = cfg.pool_width - 3.0 * cfg.dx
    compression = np.maximum.accumulate(rows[:, col["rho_max"]] / cfg.rho0 - 1.0)
    hit = np.nonzero(rows[:, col["x_max"]] >= impact_limit)[0]
    wall_impact_time = float(time[hit[0]]) if hit.size else None
    compression_at_impact = float(compression[hit[0]]) if hit.size else None
    speed = rows[:, col["max_speed"]]
    speed_peak = float(speed.max())
    sampled_speed = max(float(speed[k]) for k in sample)
    peak_compression = float(compression[-1])
    if adaptive:
        step_dt = rows[:, col["step_dt"]]
        courant = float(np.max(step_dt * (cfg.sound_speed + speed) / h))
        dt_min, dt_max = float(rows[:, col["dt"]].min()), float(rows[:, col["dt"]].max())
    else:
        courant = cfg.dt * (cfg.sound_speed + speed_peak) / h
        dt_min = dt_max = cfg.dt
    return DamBreakRun(
        times=[float(time[k]) for k in sample],
        front=[float(rows[k, col["x_max"]]) for k in sample],
        kinetic_energy=[float(rows[k, col["kinetic_energy"]]) for k in sample],
        potential_energy=[float(rows[k, col["potential_energy"]]) for k in sample],
        mass=initial_mass, escaped=int(escaped), rebuilds=int(stats["rebuilds"]),
        max_speed=max(sampled_speed, speed_peak),
        max_density_compression=peak_compression,
        compression_before_wall_impact=(peak_compression if compression_at_impact is None
                                        else compression_at_impact),
        wall_impact_time=wall_impact_time, initial_mass=initial_mass,
        final_mass=float(mass.sum()), final_positions=final_pos, steps_taken=steps_taken,
        simulated_time=float(time[steps_taken]), dt_min=dt_min, dt_max=dt_max,
        max_acoustic_courant=courant, snapshots=snapshots, max_tvc_shift=float(tvc_peak),
        max_tvc_clipped=float(tvc_clip_peak))
