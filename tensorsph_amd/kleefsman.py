"""3-D dam break against a box obstacle (kleefsman_obstacle_3d.py:553-726) on
the resident stepper.

``kleefsman_state`` builds the reference's particle sets (column, tank walls
and the box, kleefsman_obstacle_3d.py:570-577); ``run_case`` advances them
with rhs_3d physics, the Verlet pair cache, symplectic Euler,
apply_boundary_3d and clamp_out_of_obstacle, all resident on the GPU, and
returns the per-step force and moment on the box, its impulse and the clamp
bookkeeping.  The water-height gauges, pressure taps and checkpoint files of
the reference's driver are outside the particle pipeline and not built.
"""
from __future__ import annotations

import numpy as np

from .cases import make_column_3d, mass_3d, obstacle_lattice_3d, wall_lattice_3d
from .config import Kleefsman3DConfig
from .engine import WcsphSimulation


def dx_for_fluid(n_fluid: float) -> float:
    """Lattice spacing giving about n_fluid particles in the 1.228 x 0.55 x 1.0 m column."""
    return (1.228 * 0.55 * 1.0 / n_fluid) ** (1.0 / 3.0)


def kleefsman_state(cfg: Kleefsman3DConfig):
    """(fluid pos, fluid mass, wall pos, obstacle rows, box centre, box lo / hi)."""
    pos, mass = make_column_3d(cfg)
    tank = wall_lattice_3d(cfg)
    box = obstacle_lattice_3d(cfg)
    wall = np.concatenate([tank, box])
    rows = (tank.shape[0], wall.shape[0])
    centre = np.array([cfg.obstacle_x + cfg.obstacle_length / 2, cfg.obstacle_height / 2,
                       cfg.obstacle_z + cfg.obstacle_width / 2])
    lo = np.array([cfg.obstacle_x, 0.0, cfg.obstacle_z])
    hi = np.array([cfg.obstacle_x1(), cfg.obstacle_height, cfg.obstacle_z1()])
    inside = np.all((pos > lo) & (pos < hi), axis=1)
    if inside.any():
        raise ValueError(f"{int(inside.sum())} initial fluid particles sit inside the obstacle")
    return pos, mass, wall, rows, centre, (lo, hi)


def make_simulation(cfg: Kleefsman3DConfig, precision: str = "fp32", device="cuda"):
    pos, mass, wall, rows, centre, box = kleefsman_state(cfg)
    h = cfg.h_factor * cfg.dx
    wall_mass = mass_3d(cfg.dx, cfg.h_factor, cfg.rho0, cfg.density_calibration)
    sim = WcsphSimulation(pos, np.zeros_like(pos), mass, wall, h=h, rho0=cfg.rho0,
                          sound_speed=cfg.sound_speed, gamma=cfg.gamma, gravity=cfg.gravity,
                          viscosity=cfg.viscosity, dt=cfg.dt, wall_mass=wall_mass,
                          skin=cfg.pair_skin_factor * h, lo=[0.0, 0.0, 0.0],
                          hi=[cfg.pool_width, cfg.tank_top(), cfg.tank_depth], tol=0.25 * cfg.dx,
                          wall_friction=cfg.wall_friction, precision=precision, device=device,
                          obstacle={"rows": rows, "centre": centre, "box": box, "eps": 0.01 * cfg.dx})
    return sim, pos.shape[0], wall.shape[0]


def run_case(cfg: Kleefsman3DConfig, steps: int | None = None, precision: str = "fp32",
             device="cuda", meter_every: int = 0) -> dict:
    """Integrate ``steps`` steps (default: to cfg.t_end) and return the box loads."""
    steps = cfg.steps_for_t_end() if steps is None else int(steps)
    sim, nf, nw = make_simulation(cfg, precision, device)
    try:
        sim.step(steps, meter_every=meter_every)
        loads = sim.loads()
        stats = sim.stats()
        meters = sim.meters() if meter_every else None
    finally:
        sim.close()
    force, moment = loads[:, 1:4], loads[:, 4:7]
    return {"fluid_particles": nf, "wall_particles": nw, "steps": steps, "dt": cfg.dt,
            "obstacle_force_N": force, "obstacle_moment_Nm": moment,
            "obstacle_sph_impulse_Ns": (force * cfg.dt).sum(0),
            "obstacle_clamp_impulse_Ns": loads[-1, 7:10] if len(loads) else np.zeros(3),
            "obstacle_clamps": int(loads[-1, 10]) if len(loads) else 0,
            "pair_rebuilds": stats["rebuilds"], "boundary_escapes": stats["escaped"],
            "meters": meters}
