"""GPU parity of the 3-D resident stepper with a box obstacle
(kleefsman_obstacle_3d.py:553-726): loads, moment and the obstacle clamp.

Bars: FP64 positions within 1e-10 of the oracle loop after 40 steps, the box
force and moment of every step within 1e-9 of the largest oracle value, the
same list builds; the clamp moves the same particles to the same places with
the same impulse; FP32-compensated loads within 1e-4 of FP64.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_util import n, scaled_err

pytestmark = pytest.mark.gpu


def _cfg(**kw):
    from tensorsph_amd.config import Kleefsman3DConfig
    base = dict(dx=0.05, obstacle_x=1.25, sound_speed=20.0, dt=2.0e-4, viscosity=1e-6)
    base.update(kw)
    return Kleefsman3DConfig(**base)


def _oracle(cfg, steps):
    from tensorsph_amd.cases import mass_3d
    from tensorsph_amd.kleefsman import kleefsman_state
    pos, mass, wall, rows, centre, box = kleefsman_state(cfg)
    h = cfg.h_factor * cfg.dx
    st = orc.Stepper3D(pos, np.zeros_like(pos), mass, wall, h=h, rho0=cfg.rho0, c0=cfg.sound_speed,
                       gamma=cfg.gamma, gravity=cfg.gravity, nu=cfg.viscosity, dt=cfg.dt,
                       wall_mass=mass_3d(cfg.dx, cfg.h_factor, cfg.rho0, cfg.density_calibration),
                       skin=cfg.pair_skin_factor * h, lo=[0, 0, 0],
                       hi=[cfg.pool_width, cfg.tank_top(), cfg.tank_depth], tol=0.25 * cfg.dx,
                       obstacle_rows=rows, centre=centre, box=box, eps=0.01 * cfg.dx,
                       wall_friction=cfg.wall_friction)
    for _ in range(steps):
        st.step()
    return st


def test_obstacle_loads_track_oracle():
    from tensorsph_amd.kleefsman import make_simulation
    cfg = _cfg()
    steps = 40
    sim, nf, nw = make_simulation(cfg, "fp64")
    sim.step(steps)
    loads = sim.loads()
    state = sim.read()
    stats = sim.stats()
    sim.close()
    ref = _oracle(cfg, steps)
    assert loads.shape == (steps, 12)
    assert np.all(loads[:, 11] == cfg.dt)          # the step size column
    assert np.array_equal(loads[:, 0], np.arange(steps))
    ref_loads = np.array(ref.loads)
    scale = np.abs(ref_loads[:, :3]).max()
    assert scale > 0.0
    assert np.abs(loads[:, 1:7] - ref_loads).max() < 1e-9 * max(scale, np.abs(ref_loads[:, 3:]).max())
    assert np.abs(n(state["position"]) - ref.pos).max() < 1e-10
    assert stats["rebuilds"] == ref.rebuilds


def test_obstacle_clamp_matches_oracle():
    """A box laid over part of the water column: the particles inside it are
    pushed out through their nearest face at the first step."""
    from tensorsph_amd.cases import make_column_3d, mass_3d, obstacle_lattice_3d, wall_lattice_3d
    from tensorsph_amd.engine import WcsphSimulation
    cfg = _cfg(obstacle_x=0.52)                      # overlaps the column (x < 1.228)
    pos, mass = make_column_3d(cfg)
    tank = wall_lattice_3d(cfg)
    wall = tank                                      # clamp only: no obstacle stations
    lo = np.array([cfg.obstacle_x, 0.0, cfg.obstacle_z])
    hi = np.array([cfg.obstacle_x1(), cfg.obstacle_height, cfg.obstacle_z1()])
    inside = int(np.all((pos > lo) & (pos < hi), axis=1).sum())
    assert inside > 0
    h = cfg.h_factor * cfg.dx
    wm = mass_3d(cfg.dx, cfg.h_factor, cfg.rho0, cfg.density_calibration)
    kw = dict(h=h, rho0=cfg.rho0, gamma=cfg.gamma, gravity=cfg.gravity, dt=cfg.dt, wall_mass=wm,
              skin=cfg.pair_skin_factor * h, lo=[0, 0, 0], hi=[cfg.pool_width, cfg.tank_top(), cfg.tank_depth],
              tol=0.25 * cfg.dx)
    sim = WcsphSimulation(pos, np.zeros_like(pos), mass, wall, sound_speed=cfg.sound_speed,
                          viscosity=cfg.viscosity, precision="fp64",
                          obstacle={"rows": (0, 0), "centre": (0, 0, 0), "box": (lo, hi), "eps": 0.01 * cfg.dx},
                          **kw)
    sim.step(3)
    got = n(sim.read()["position"])
    loads = sim.loads()
    sim.close()
    st = orc.Stepper3D(pos, np.zeros_like(pos), mass, wall, c0=cfg.sound_speed, nu=cfg.viscosity,
                       obstacle_rows=(0, 0), centre=(0, 0, 0), box=(lo, hi), eps=0.01 * cfg.dx, **kw)
    for _ in range(3):
        st.step()
    assert st.clamped >= inside
    assert np.abs(got - st.pos).max() < 1e-10
    assert not np.all((got > lo) & (got < hi), axis=1).any()
    # cumulative clamp bookkeeping after the last step
    assert loads.shape == (3, 12)
    assert int(loads[-1, 10]) == st.clamped
    imp = np.asarray(st.clamp_impulse)
    assert np.abs(loads[-1, 7:10] - imp).max() <= 1e-9 * max(np.abs(imp).max(), 1e-30)


def test_obstacle_loads_fp32_compensated():
    from tensorsph_amd.kleefsman import make_simulation
    cfg = _cfg()
    out = {}
    for prec in ("fp32", "fp64"):
        sim, _, _ = make_simulation(cfg, prec)
        sim.step(20)
        out[prec] = sim.loads()
        sim.close()
    a, b = out["fp32"][:, 1:7], out["fp64"][:, 1:7]
    assert scaled_err(a[:, :3], b[:, :3]) < 1e-4
    assert scaled_err(a[:, 3:], b[:, 3:]) < 1e-4
