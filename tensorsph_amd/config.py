"""Case and solver parameters, field for field the reference's dataclasses.

``Config`` mirrors droplet_impact.Config (droplet_impact.py:55-269);
``DamBreakConfig`` dam_break.py:45-120; ``DamBreak3DConfig``
dam_break_3d.py:103-143; ``Kleefsman3DConfig`` kleefsman_obstacle_3d.py:80-126.
Same names, same defaults, so a caller's existing ``Config(...)`` works
unchanged.  Options whose operators are not on the particle pipeline
(surface tension, compression regularisation, the transport-velocity
formulation, initial-mass calibration) are accepted as fields but rejected
by ``validate_config`` when switched on, with a message naming the option.
The surface-pressure closure, Shepard renormalisation and boundary-consistent
density run in the operator rhs and the resident stepper; Lind shifting and
the transport-velocity correction (``shift_every`` / ``tvc_every``) run in
``dam_break.run_case``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

VISCOUS_LAPLACIAN_FACTOR = 2.0   # Morris factor 2, droplet_impact.py:48


class SimulationUnstableError(FloatingPointError):
    """A run violated one of its numerical acceptance criteria (droplet_impact.py:51)."""


@dataclass
class Config:
    dx: float = 0.025
    h_factor: float = 1.3
    rho0: float = 1000.0
    sound_speed: float = 20.0
    gamma: float = 7.0
    gravity: float = 9.81
    viscosity: float = 0.01
    dt: float = 4.0e-4
    steps: int = 750
    output_every: int = 50
    pool_width: float = 1.0
    pool_depth: float = 0.35
    drop_radius: float = 0.10
    drop_center_y: float = 0.49
    device: str = "auto"
    boundary: str = "wall"
    wall_layers: int = 3
    max_courant: float = 0.3
    max_density_error: float = 0.05
    max_mach: float = 0.10
    fail_fast: bool = True
    pair_skin_factor: float = 0.25
    hydrostatic_init: bool = False
    calibrate_initial_mass: bool = False
    wall_friction: float = 0.0
    density_diffusion: float = 0.0
    density_diffusion_form: str = "molteni"
    density_diffusion_eigenvalue_floor: float = 0.1
    density_calibration: bool = True
    adaptive_dt: bool = False
    dt_growth: float = 1.1
    dt_safety: float = 0.9
    surface_tension_coefficient: float = 0.0
    contact_angle: float | None = None
    contact_angle_reach: float = 2.0
    contact_angle_model: str = "rotate"
    csf_scheme: str = "consistent"
    csf_delta_normalization: str = "legacy"
    csf_curvature: str = "fit_quadratic"
    surface_pressure_closure: bool = False
    surface_pressure_threshold: float = 0.99
    surface_pressure_donors: str = "interior"
    surface_pressure_layers: int = 2
    density_renormalization: bool = False
    boundary_consistent_density: bool = False
    density_regularization: str = "none"
    compression_regularization_rule: str = "free_surface"
    compression_normalization: str = "reference_mean"
    transport_velocity: str = "off"
    transport_velocity_strength: float = 1.0
    transport_velocity_v_ref: float | None = None
    # Harness options of the impact benchmark (not reference fields): the
    # drop's initial velocity and the wall height.  Their defaults reproduce
    # the reference run (drop at rest, walls up to ``domain_top``).
    initial_velocity: tuple = (0.0, 0.0)
    wall_top: float | None = None


@dataclass
class DamBreakConfig(Config):
    dx: float = 0.01
    tank_height: float = 0.75
    column_width: float = 0.2
    column_height: float = 0.4
    pool_width: float = 2.0
    pool_depth: float = 0.75
    dt: float = 1.0e-4
    steps: int = 6000
    output_every: int = 100
    device: str = "cpu"
    hydrostatic_init: bool = True
    shift_every: int = 0
    shift_strength: float = 1.0
    shift_limit: float = 0.25
    shift_v_ref: float | None = None
    tvc_every: int = 0
    tvc_alpha: float | None = None
    tvc_scope: str = "all"
    tvc_bulk_c: float = 0.95
    tvc_bulk_g: float = 0.02
    tvc_max_shift: float | None = None
    tvc_limiter: str | None = None
    tvc_limiter_slope: float = 100.0

    def tank_top(self) -> float:  # type: ignore[override]
        return float(self.tank_height)

    def dt_from_courant(self) -> float:
        return self.max_courant * self.h_factor * self.dx / self.sound_speed


@dataclass
class DamBreak3DConfig(Config):
    dx: float = 0.02
    tank_height: float = 0.6
    tank_depth: float = 0.3
    column_width: float = 0.15
    column_depth: float = 0.15
    column_height: float = 0.3
    pool_width: float = 1.0
    pool_depth: float = 0.6
    dt: float = 5.0e-5
    steps: int = 4000
    output_every: int = 100
    device: str = "cpu"
    hydrostatic_init: bool = True
    tvc_every: int = 0
    tvc_alpha: float | None = None
    tvc_scope_tol: float | None = None
    tvc_max_shift: float | None = None

    def tank_top(self) -> float:  # type: ignore[override]
        return float(self.tank_height)

    def dt_from_courant(self) -> float:
        return self.max_courant * self.h_factor * self.dx / self.sound_speed


TANK_LENGTH_M = 3.22
TANK_HEIGHT_M = 1.00
TANK_WIDTH_M = 1.00
COLUMN_LENGTH_M = 1.228
COLUMN_HEIGHT_M = 0.550
OBSTACLE_LENGTH_M = 0.161
OBSTACLE_HEIGHT_M = 0.161
OBSTACLE_WIDTH_M = 0.403
OBSTACLE_X_M = TANK_LENGTH_M - 0.744 - OBSTACLE_LENGTH_M / 2
H2_X_M = TANK_LENGTH_M - 0.992
# the gauge protocol of the 2-D runs, kept for the 3-D gauges
FROZEN_APERTURE_M = 0.08
FROZEN_STATION_OFFSET_M = OBSTACLE_X_M - H2_X_M
FROZEN_ARRIVAL_THRESHOLD_M = 0.02


@dataclass
class Kleefsman3DConfig(DamBreak3DConfig):
    dx: float = 0.04
    pool_width: float = TANK_LENGTH_M
    tank_depth: float = TANK_WIDTH_M
    tank_height: float = TANK_HEIGHT_M
    column_width: float = COLUMN_LENGTH_M
    column_height: float = COLUMN_HEIGHT_M
    column_depth: float = TANK_WIDTH_M
    obstacle_x: float = OBSTACLE_X_M
    obstacle_length: float = OBSTACLE_LENGTH_M
    obstacle_height: float = OBSTACLE_HEIGHT_M
    obstacle_width: float = OBSTACLE_WIDTH_M
    obstacle_z: float = (TANK_WIDTH_M - OBSTACLE_WIDTH_M) / 2
    station_offset: float = FROZEN_STATION_OFFSET_M
    gauge_aperture: float = FROZEN_APERTURE_M
    arrival_threshold: float = FROZEN_ARRIVAL_THRESHOLD_M
    pressure_disk_diameter: float = 0.0
    dt: float = 2.6e-4
    t_end: float = 2.6
    output_every: int = 25
    max_steps: int = 0
    wall_layers: int = 2
    pair_skin_factor: float = 0.25
    # obstacle_lattice_3d fills the box (kleefsman_obstacle_3d.py:187-205);
    # "shell" keeps only the outer ``wall_layers`` of it.
    obstacle_fill: str = "filled"

    def obstacle_x1(self) -> float:
        return self.obstacle_x + self.obstacle_length

    def obstacle_z1(self) -> float:
        return self.obstacle_z + self.obstacle_width

    def station_x(self) -> float:
        return self.obstacle_x - self.station_offset

    def blocked_fraction(self) -> float:
        return self.obstacle_width / self.tank_depth

    def steps_for_t_end(self) -> int:
        return int(math.ceil(self.t_end / self.dt)) + 2


# --------------------------------------------------------------------------
# Step-size limits, droplet_impact.py:495-552 and :1389-1421
# --------------------------------------------------------------------------

def acoustic_courant(dt: float, cfg: Config, max_speed: float = 0.0) -> float:
    return dt * (cfg.sound_speed + max_speed) / (cfg.h_factor * cfg.dx)


def suggested_dt(cfg: Config, max_courant: float | None = None) -> float:
    limit = cfg.max_courant if max_courant is None else max_courant
    return limit * cfg.h_factor * cfg.dx / cfg.sound_speed


def acoustic_dt_limit(cfg: Config, max_speed: float = 0.0) -> float:
    return cfg.max_courant * cfg.h_factor * cfg.dx / (cfg.sound_speed + max(0.0, max_speed))


def viscous_dt_limit(cfg: Config, safety: float = 0.125) -> float:
    h = cfg.h_factor * cfg.dx
    if cfg.viscosity <= 0.0:
        return float("inf")
    return safety * h * h / cfg.viscosity


def diffusion_dt_limit(cfg: Config, safety: float = 0.125) -> float:
    if cfg.density_diffusion <= 0.0:
        return float("inf")
    h = cfg.h_factor * cfg.dx
    return safety * h / (cfg.density_diffusion * cfg.sound_speed)


def gravity_dt_limit(cfg: Config, safety: float = 0.25) -> float:
    if cfg.gravity <= 0.0:
        return float("inf")
    return safety * math.sqrt(cfg.h_factor * cfg.dx / cfg.gravity)


def stable_dt(cfg: Config, max_speed: float = 0.0) -> float:
    return min(acoustic_dt_limit(cfg, max_speed), viscous_dt_limit(cfg),
               gravity_dt_limit(cfg), diffusion_dt_limit(cfg))


def next_dt(cfg: Config, dt: float, max_speed: float = 0.0) -> float:
    target = cfg.dt_safety * stable_dt(cfg, max_speed)
    grown = dt * cfg.dt_growth if cfg.dt_growth >= 1.0 else target
    return min(target, grown, cfg.dt)


def adaptive_controls(cfg: Config) -> dict:
    """The adaptive-step rule of next_dt (droplet_impact.py:495-552) for the
    resident stepper: the run integrates to cfg.steps * cfg.dt; the acoustic
    limit's numerator max_courant * h_factor * dx is formed as the reference
    forms it, and the speed-independent limits are folded into one."""
    return {"target_time": cfg.steps * cfg.dt,
            "courant_num": cfg.max_courant * cfg.h_factor * cfg.dx,
            "dt_limit": min(viscous_dt_limit(cfg), gravity_dt_limit(cfg), diffusion_dt_limit(cfg)),
            "dt_safety": cfg.dt_safety, "dt_growth": cfg.dt_growth}


_UNSUPPORTED = (
    ("surface_tension_coefficient", lambda c: c.surface_tension_coefficient > 0.0),
    ("surface_pressure_closure", lambda c: bool(c.surface_pressure_closure)),
    ("density_renormalization", lambda c: bool(c.density_renormalization)),
    ("boundary_consistent_density", lambda c: bool(c.boundary_consistent_density)),
    ("density_regularization", lambda c: c.density_regularization != "none"),
    ("transport_velocity", lambda c: c.transport_velocity != "off"),
    ("calibrate_initial_mass", lambda c: bool(c.calibrate_initial_mass)),
)


# branches the operator rhs (tsph_wcsph_rhs_ex) and the resident stepper
# (queue_wall_stage, the Shepard density sweep) both implement
_RHS_ONLY = ("surface_pressure_closure", "density_renormalization", "boundary_consistent_density")


def check_supported(cfg: Config, rhs: bool = False, shifting: bool = False) -> None:
    """Reject opt-in branches that are not part of the particle pipeline
    (with rhs=True the closure, Shepard and boundary-consistent branches pass,
    as the rhs and the stepper run them; with shifting=True the caller runs
    Lind shifting and the transport-velocity correction, shift_every /
    tvc_every)."""
    for name, on in _UNSUPPORTED:
        if rhs and name in _RHS_ONLY:
            continue
        if on(cfg):
            raise NotImplementedError(
                f"Config.{name} is an opt-in rhs branch outside the GPU particle "
                f"pipeline; switch it off to run on tensorsph_amd")
    for name in ("shift_every", "tvc_every"):
        if shifting:
            continue
        if getattr(cfg, name, 0):
            raise NotImplementedError(
                f"{type(cfg).__name__}.{name} (position shifting) is not part of the "
                f"GPU particle pipeline")


def validate_config(cfg: Config) -> None:
    """droplet_impact.validate_config (droplet_impact.py:1319-1386)."""
    if cfg.dx <= 0 or cfg.dt <= 0 or cfg.h_factor <= 0:
        raise ValueError("dx, dt, and h_factor must be positive")
    if cfg.rho0 <= 0 or cfg.sound_speed <= 0 or cfg.gamma <= 0:
        raise ValueError("rho0, sound_speed, and gamma must be positive")
    if cfg.steps < 0 or cfg.output_every <= 0:
        raise ValueError("steps must be non-negative and output_every must be positive")
    if cfg.pool_width <= 0 or cfg.pool_depth <= 0 or cfg.drop_radius <= 0:
        raise ValueError("pool dimensions and drop_radius must be positive")
    if cfg.boundary not in ("wall", "clamp"):
        raise ValueError("boundary must be 'wall' or 'clamp'")
    if not 0.0 < cfg.max_courant <= 1.0:
        raise ValueError("max_courant must be in (0, 1]")
    if cfg.max_density_error <= 0 or cfg.max_mach <= 0:
        raise ValueError("max_density_error and max_mach must be positive")
    if cfg.pair_skin_factor < 0:
        raise ValueError("pair_skin_factor must be non-negative")
    if not 0.0 <= cfg.wall_friction <= 1.0:
        raise ValueError("wall_friction must be in [0, 1]")
    if cfg.density_diffusion < 0.0:
        raise ValueError("density_diffusion must be non-negative")
    if cfg.density_diffusion_form not in ("molteni", "antuono"):
        raise ValueError("density_diffusion_form must be one of ('molteni', 'antuono'), "
                         f"got {cfg.density_diffusion_form!r}")
    if not 0.0 <= cfg.density_diffusion_eigenvalue_floor < 1.0:
        raise ValueError("density_diffusion_eigenvalue_floor must be in [0, 1)")
    if not math.isfinite(cfg.dt_growth) or cfg.dt_growth < 1.0:
        raise ValueError("dt_growth must be finite and at least 1.0")
    if not 0.0 < cfg.dt_safety <= 1.0:
        raise ValueError("dt_safety must be in (0, 1]")
    courant = acoustic_courant(cfg.dt, cfg)
    if courant > cfg.max_courant * (1.0 + 1.0e-9):
        raise ValueError(f"dt={cfg.dt:g} gives acoustic Courant {courant:.3f} > max_courant "
                         f"{cfg.max_courant:g}; use dt <= {suggested_dt(cfg):.3e} s")
    lim = viscous_dt_limit(cfg)
    if cfg.dt > lim * (1.0 + 1.0e-9):
        raise ValueError(f"dt={cfg.dt:g} exceeds the viscous limit {lim:.3e} s "
                         f"(viscosity={cfg.viscosity:g}, h={cfg.h_factor * cfg.dx:g}); the "
                         f"explicit Laplacian is unstable above it")
    lim = gravity_dt_limit(cfg)
    if cfg.dt > lim * (1.0 + 1.0e-9):
        raise ValueError(f"dt={cfg.dt:g} exceeds the gravity limit {lim:.3e} s")
    lim = diffusion_dt_limit(cfg)
    if cfg.dt > lim * (1.0 + 1.0e-9):
        raise ValueError(f"dt={cfg.dt:g} exceeds the density-diffusion limit {lim:.3e} s "
                         f"(delta={cfg.density_diffusion:g})")
    check_supported(cfg, rhs=True, shifting=True)
