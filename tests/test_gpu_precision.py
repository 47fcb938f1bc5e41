"""FP32-compensated resident steps against FP64 on a deep water column.

At depth H a row's q = p / rho^2 is about H / h times the pressure contrast
it carries to a neighbour, so rounding q (or the symmetric pressure term
m_j (q_i + q_j) grad W_ij) to FP32 puts about (H / h) 2^-24 of the pressure
scale into the force.  The FP32 steps therefore sum the pressure terms as
-2 q_i G_i (G_i = sum_j m_j grad W_ij, FP64 in the density sweep) plus the
contrast terms -m_j (q_j - q_i) grad W_ij, with q kept as two floats.  The
same amplification applies to the masses: a density summed from masses
rounded to FP32 moves p by c0^2 rho0 2^-24 / sqrt(n) against contrasts of
rho g dx, so the FP32 steps keep each mass as two floats as well (the
printed "vs FP64 on FP32 masses" figure is the error those roundings alone
would leave, about 1.4e-5 here).

Workload: the 1 x 2 m column in the 4 x 3 m tank of the dam-break bench at
200k fluid particles (H / h ~ 490), one step in each precision from the same
state: the hydrostatic start and the collapse after 3,000 FP64 steps.
Bars (north star: FP32 forces within 1e-5 of FP64): acceleration within
1e-5 of the FP64 field's largest magnitude, density within 2e-7 relative
(the FP32 storage of rho), on every layout of the 2-D stepper.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dam_break(n_target):
    from tensorsph_amd.cases import lattice_mass_unit, make_column, wall_lattice
    from tensorsph_amd.config import DamBreakConfig
    dx = math.sqrt(2.0 / n_target)
    c0 = 10.0 * math.sqrt(2.0 * 9.81 * 2.0)
    h = 1.3 * dx
    cfg = DamBreakConfig(dx=dx, tank_height=3.0, column_width=1.0, column_height=2.0,
                         pool_width=4.0, sound_speed=c0, dt=0.3 * h / c0, viscosity=0.01)
    pos, mass = make_column(cfg)
    wall = wall_lattice(cfg, top=cfg.tank_top())
    return cfg, pos, mass, wall, lattice_mass_unit(cfg)


def _sim(cfg, pos, vel, mass, wall, wmass, precision, layout):
    from tensorsph_amd.engine import WcsphSimulation
    h = cfg.h_factor * cfg.dx
    sim = WcsphSimulation(pos, vel, mass, wall, h=h, rho0=cfg.rho0, sound_speed=cfg.sound_speed,
                          gamma=cfg.gamma, gravity=cfg.gravity, viscosity=cfg.viscosity, dt=cfg.dt,
                          wall_mass=wmass, skin=cfg.pair_skin_factor * h, lo=[0.0, 0.0],
                          hi=[cfg.pool_width, cfg.tank_top()], tol=0.25 * cfg.dx, precision=precision)
    if layout in ("lists", "halo"):
        sim.configure(tiles=False, halo=layout == "halo")
    else:
        sim.configure(tiles=True, tile=layout)
    return sim


def _scaled(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def column():
    cfg, pos, mass, wall, wmass = _dam_break(2.0e5)
    base = _sim(cfg, pos, np.zeros_like(pos), mass, wall, wmass, "fp64", "halo")
    try:
        base.step(3000)
        st = base.read()
        developed = (st["position"].double().cpu().numpy(), st["velocity"].double().cpu().numpy())
    finally:
        base.close()
    return cfg, mass, wall, wmass, {"hydrostatic": (pos, np.zeros_like(pos)), "developed": developed}


@pytest.mark.parametrize("state", ["hydrostatic", "developed"])
@pytest.mark.parametrize("layout", ["halo", "lists", (8, 4)])
def test_fp32c_force_on_deep_column_within_1e5_of_fp64(column, state, layout):
    cfg, mass, wall, wmass, states = column
    pos, vel = states[state]
    print(f"deep column: state {state}, layout {layout}, {pos.shape[0]} fluid", flush=True)
    out = {}
    m32 = mass.astype(np.float32).astype(np.float64)
    for prec, m in (("fp32", mass), ("fp64", mass), ("fp64r", m32)):
        sim = _sim(cfg, pos, vel, m, wall, wmass, prec[:4], layout)
        try:
            sim.step(1)
            r = sim.read()
            out[prec] = {k: r[k].double().cpu().numpy() for k in ("density", "pressure", "acceleration")}
        finally:
            sim.close()
    acc = _scaled(out["fp32"]["acceleration"], out["fp64"]["acceleration"])
    accr = _scaled(out["fp32"]["acceleration"], out["fp64r"]["acceleration"])
    rho = float(np.abs(out["fp32"]["density"] / out["fp64"]["density"] - 1.0).max())
    a64 = out["fp64"]["acceleration"]
    mag = np.sqrt((a64 ** 2).sum(1))
    e = np.sqrt(((out["fp32"]["acceleration"] - a64) ** 2).sum(1))
    k = int(np.argmax(e))
    print(f"  acceleration {acc:.3e} (vs FP64 on FP32 masses {accr:.3e}), density {rho:.3e}; "
          f"max|a| {mag.max():.4g}, p99 |a| {np.percentile(mag, 99):.4g}, worst row {k} at {pos[k]} "
          f"|a| {mag[k]:.4g} err {e[k]:.3g}; rows with err > 1e-5 max|a|: {(e > 1e-5 * mag.max()).sum()}",
          flush=True)
    assert acc < 1e-5
    assert rho < 2e-7


def test_fp32c_run_tracks_the_fp64_run_over_2000_steps():
    """The dam break of the bench at its smallest size (100k fluid), run from
    rest for 2,000 steps in each precision on the default layout, with the
    mass / momentum / energy meters every 20 steps.  Bars: the mass meter
    equal to 1e-12 relative (the FP32 run meters its masses as two floats);
    momentum, kinetic and potential energy and the density sum within 1e-6 of
    their largest magnitude over the run at every metered step (measured on
    MI355X: at most 8.4e-8); the list-build counts equal; the final positions
    within 1e-6 of the column height (measured 1.6e-8).  Per-particle final
    velocities and densities carry 2,000 steps of accumulated FP32 force
    rounding (about 1e-6 of the largest acceleration per step): bars 2e-4 of
    the largest speed and 2e-5 of the largest density (measured 7.1e-5 and
    6.9e-6)."""
    from tensorsph_amd.engine import METER_KEYS
    cfg, pos, mass, wall, wmass = _dam_break(1.0e5)
    print(f"dam break run: {pos.shape[0]} fluid + {wall.shape[0]} wall, 2000 steps per precision", flush=True)
    runs = {}
    for prec in ("fp32", "fp64"):
        sim = _sim(cfg, pos, np.zeros_like(pos), mass, wall, wmass, prec, "halo")
        try:
            sim.step(2000, meter_every=20)
            r = sim.read()
            runs[prec] = (sim.meters(), {k: r[k].double().cpu().numpy() for k in ("position", "velocity", "density")},
                          sim.stats())
        finally:
            sim.close()
    m32, s32, st32 = runs["fp32"]
    m64, s64, st64 = runs["fp64"]
    assert m32.shape == m64.shape and m32.shape[0] >= 100
    col = {k: i for i, k in enumerate(METER_KEYS)}
    assert np.array_equal(m32[:, col["step"]], m64[:, col["step"]])
    mass_err = float(np.abs(m32[:, col["mass"]] / m64[:, col["mass"]] - 1.0).max())
    errs = {}
    for k in ("momentum_x", "momentum_y", "kinetic_energy", "potential_energy", "rho_sum"):
        a, b = m32[:, col[k]], m64[:, col[k]]
        errs[k] = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    x_err = float(np.abs(s32["position"] - s64["position"]).max() / cfg.column_height)
    v_err = _scaled(s32["velocity"], s64["velocity"])
    r_err = _scaled(s32["density"], s64["density"])
    print(f"  mass {mass_err:.2e}; " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items())
          + f"; final x {x_err:.2e}, v {v_err:.2e}, rho {r_err:.2e}; rebuilds {st32['rebuilds']} / {st64['rebuilds']}",
          flush=True)
    assert mass_err < 1e-12
    for k, v in errs.items():
        assert v < 1e-6, k
    assert x_err < 1e-6
    assert v_err < 2e-4 and r_err < 2e-5
    assert st32["rebuilds"] == st64["rebuilds"]
