"""GPU parity of the resident stepper against the reference's own run() and
the FP64 oracle loop.

Bars: FP64 stepper meters (mass, momentum, kinetic energy, density extrema)
within 1e-9 relative of the reference's run() rows over 40 steps, positions
within 1e-9 m; FP32-compensated stepper within 1e-5 (scaled) of FP64 after 20
steps; two runs bitwise identical.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from tests.gpu_util import dev, n, scaled_err, t

pytestmark = pytest.mark.gpu


def _use(sim, layout):
    """"default": the library's choice; "tiles": 8 x 4 cell tiles; "lists",
    "halo"; or a tile shape."""
    if layout in ("lists", "halo"):
        sim.configure(tiles=False, halo=layout == "halo")
    elif layout == "tiles":
        sim.configure(tiles=True, tile=(8, 4))
    elif layout != "default":
        sim.configure(tiles=True, tile=layout)


def _impact_sim(precision, golden, layout="default"):
    from tensorsph_amd.cases import impact_state, domain_top
    from tensorsph_amd.engine import WcsphSimulation
    g = golden("run2d")
    dx, steps, dt, nu = g["params"]
    cfg, pos, mass, wall = impact_state(dx=dx, pool_width=0.4, pool_depth=0.15, drop_radius=0.05,
                                        drop_center_y=0.25, viscosity=nu, dt=dt)
    h = cfg.h_factor * cfg.dx
    sim = WcsphSimulation(pos, np.zeros_like(pos), mass, wall, h=h, rho0=cfg.rho0,
                          sound_speed=cfg.sound_speed, gamma=cfg.gamma, gravity=cfg.gravity,
                          viscosity=nu, dt=dt, wall_mass=float(mass[0]), skin=0.25 * h,
                          lo=[0.0, 0.0], hi=[cfg.pool_width, domain_top(cfg)],
                          tol=0.25 * cfg.dx, precision=precision)
    _use(sim, layout)
    return sim, g, int(steps), mass


@pytest.mark.parametrize("layout", ["default", "tiles", "lists", "halo", (4, 4), (16, 4)])
def test_stepper_fp64_tracks_reference_run(golden, layout):
    from tensorsph_amd.engine import METER_KEYS
    sim, g, steps, mass = _impact_sim("fp64", golden, layout)
    sim.step(steps, meter_every=1)
    rows = sim.meters()
    keys = list(g["row_keys"])
    ref = g["rows"]
    assert rows.shape[0] == steps
    col = {k: METER_KEYS.index(k) for k in METER_KEYS}
    ke_scale = ref[:, keys.index("kinetic_energy")].max()
    for k in range(steps):
        assert abs(rows[k, col["mass"]] - ref[k, keys.index("mass_total")]) < 1e-12 * ref[k, keys.index("mass_total")]
        assert abs(rows[k, col["kinetic_energy"]] - ref[k, keys.index("kinetic_energy")]) < 1e-9 * ke_scale
        assert abs(rows[k, col["rho_max"]] - ref[k, keys.index("rho_max")]) < 1e-9 * 1000
        assert abs(rows[k, col["rho_min"]] - ref[k, keys.index("rho_min")]) < 1e-9 * 1000
        assert abs(rows[k, col["rho_sum"]] / mass.size - ref[k, keys.index("rho_mean")]) < 1e-9 * 1000
        assert abs(rows[k, col["y_min"]] - ref[k, keys.index("y_min")]) < 1e-12
    st = sim.read()
    assert np.abs(n(st["position"]) - g["final_pos"]).max() < 1e-9
    stats = sim.stats()
    assert stats["steps"] == steps and stats["rebuilds"] >= 1


@pytest.mark.parametrize("layout", ["tiles", "lists", "halo"])
def test_stepper_matches_oracle_loop_on_dam_break(layout):
    from tensorsph_amd.config import DamBreakConfig
    from tensorsph_amd.cases import make_column, wall_lattice, lattice_mass_unit
    from tensorsph_amd.engine import WcsphSimulation
    cfg = DamBreakConfig(dx=0.02, tank_height=0.6, column_width=0.3, column_height=0.4,
                         pool_width=0.8, sound_speed=20.0, dt=3.0e-4, viscosity=1e-3)
    pos, mass = make_column(cfg)
    wall = wall_lattice(cfg, top=cfg.tank_top())
    h = cfg.h_factor * cfg.dx
    kw = dict(h=h, rho0=cfg.rho0, gamma=cfg.gamma, gravity=cfg.gravity, dt=cfg.dt,
              wall_mass=lattice_mass_unit(cfg), skin=0.25 * h)
    sim = WcsphSimulation(pos, np.zeros_like(pos), mass, wall, sound_speed=cfg.sound_speed,
                          viscosity=cfg.viscosity, lo=[0, 0], hi=[cfg.pool_width, cfg.tank_top()],
                          tol=0.25 * cfg.dx, precision="fp64", **kw)
    _use(sim, layout)
    ref = orc.Stepper2D(pos, np.zeros_like(pos), mass, wall, c0=cfg.sound_speed, nu=cfg.viscosity,
                        lo=[0, 0], hi=[cfg.pool_width, cfg.tank_top()], tol=0.25 * cfg.dx, **kw)
    steps = 60
    for _ in range(steps):
        ref.step()
    sim.step(steps)
    st = sim.read()
    assert np.abs(n(st["position"]) - ref.pos).max() < 1e-10
    assert scaled_err(n(st["velocity"]), ref.vel) < 1e-9
    assert sim.stats()["rebuilds"] == ref.rebuilds


@pytest.mark.parametrize("layout", ["tiles", "lists", "halo"])
def test_stepper_fp32_compensated_vs_fp64_and_determinism(golden, layout):
    a, g, steps, _ = _impact_sim("fp32", golden, layout)
    b, _, _, _ = _impact_sim("fp32", golden, layout)
    c, _, _, _ = _impact_sim("fp64", golden, layout)
    for s in (a, b, c):
        s.step(20)
    ra, rb, rc = a.read(), b.read(), c.read()
    for k in ra:
        assert torch.equal(ra[k], rb[k]), k
    assert scaled_err(n(ra["velocity"]), n(rc["velocity"])) < 1e-5
    assert scaled_err(n(ra["density"]), n(rc["density"])) < 1e-5
    assert np.abs(n(ra["position"]) - n(rc["position"])).max() < 1e-7


def test_tiles_and_lists_agree_on_dam_break_fp32():
    """Both sweep layouts visit the same pairs; in FP32 they may differ only by
    summation order (within 1e-5 of each other after 30 steps)."""
    from tensorsph_amd.config import DamBreakConfig
    from tensorsph_amd.cases import make_column, wall_lattice, lattice_mass_unit
    from tensorsph_amd.engine import WcsphSimulation
    cfg = DamBreakConfig(dx=0.01, tank_height=0.6, column_width=0.3, column_height=0.4,
                         pool_width=0.8, sound_speed=20.0, dt=1.5e-4, viscosity=1e-3)
    pos, mass = make_column(cfg)
    wall = wall_lattice(cfg, top=cfg.tank_top())
    h = cfg.h_factor * cfg.dx
    out = []
    for layout in ("tiles", "lists"):
        sim = WcsphSimulation(pos, np.zeros_like(pos), mass, wall, h=h, rho0=cfg.rho0,
                              sound_speed=cfg.sound_speed, gamma=cfg.gamma, gravity=cfg.gravity,
                              viscosity=cfg.viscosity, dt=cfg.dt, wall_mass=lattice_mass_unit(cfg),
                              skin=0.25 * h, lo=[0, 0], hi=[cfg.pool_width, cfg.tank_top()],
                              tol=0.25 * cfg.dx, precision="fp32")
        _use(sim, layout)
        sim.step(30)
        out.append(sim.read())
    assert scaled_err(n(out[0]["velocity"]), n(out[1]["velocity"])) < 1e-5
    assert np.abs(n(out[0]["position"]) - n(out[1]["position"])).max() < 1e-8


@pytest.mark.parametrize("layout", ["tiles", "lists", "halo"])
def test_pair_count_matches_oracle(golden, layout):
    """Regression: the pair count walks the structure of the active layout."""
    sim, g, steps, mass = _impact_sim("fp64", golden, layout)
    sim.step(3)
    st = sim.read()
    from tensorsph_amd.cases import impact_state
    dx = float(g["params"][0])
    cfg, pos, _, wall = impact_state(dx=dx, pool_width=0.4, pool_depth=0.15, drop_radius=0.05,
                                     drop_center_y=0.25)
    h = cfg.h_factor * cfg.dx
    allp = np.concatenate([n(st["position"]), wall])
    pairs = orc.cell_linked_pairs(allp, 2 * h)
    nf = pos.shape[0]
    expected = int(((pairs[:, 0] < nf) | (pairs[:, 1] < nf)).sum())
    assert sim.count_pairs(2 * h) == expected


@pytest.mark.parametrize("form", ["molteni", "antuono"])
def test_stepper_with_density_diffusion_matches_oracle_loop(form):
    """δ-diffusion (density_diffusion.py:78-197) in the resident 2-D loop runs on
    the slot-list sweeps, by default with LDS halos (the LDS cell tiles carry
    no diffusion term and refuse it): FP64 positions within 1e-10 m of the
    oracle loop after 40 steps."""
    from tensorsph_amd.config import DamBreakConfig
    from tensorsph_amd.cases import make_column, wall_lattice, lattice_mass_unit
    from tensorsph_amd.engine import WcsphSimulation
    cfg = DamBreakConfig(dx=0.02, tank_height=0.6, column_width=0.3, column_height=0.4,
                         pool_width=0.8, sound_speed=20.0, dt=3.0e-4, viscosity=1e-3)
    pos, mass = make_column(cfg)
    wall = wall_lattice(cfg, top=cfg.tank_top())
    h = cfg.h_factor * cfg.dx
    delta = 0.1
    kw = dict(h=h, rho0=cfg.rho0, gamma=cfg.gamma, gravity=cfg.gravity, dt=cfg.dt,
              wall_mass=lattice_mass_unit(cfg), skin=0.25 * h)
    sim = WcsphSimulation(pos, np.zeros_like(pos), mass, wall, sound_speed=cfg.sound_speed,
                          viscosity=cfg.viscosity, lo=[0, 0], hi=[cfg.pool_width, cfg.tank_top()],
                          tol=0.25 * cfg.dx, precision="fp64", density_diffusion=delta,
                          diffusion_form=form, **kw)
    with pytest.raises(ValueError):
        sim.configure(tiles=True)
    ref = orc.Stepper2D(pos, np.zeros_like(pos), mass, wall, c0=cfg.sound_speed, nu=cfg.viscosity,
                        lo=[0, 0], hi=[cfg.pool_width, cfg.tank_top()], tol=0.25 * cfg.dx,
                        diff_scale=2.0 * delta * cfg.sound_speed * h,
                        diff_form=1 if form == "antuono" else 0, **kw)
    steps = 40
    for _ in range(steps):
        ref.step()
    sim.step(steps)
    st = sim.read()
    assert np.abs(n(st["position"]) - ref.pos).max() < 1e-10
    assert scaled_err(n(st["velocity"]), ref.vel) < 1e-9
    assert sim.stats()["rebuilds"] == ref.rebuilds
