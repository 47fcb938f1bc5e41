"""Write the golden vectors that pin the oracle to the reference.

Run in the build container only (the reference is not on the GPU box):

    python tests/golden/make_golden.py [/path/to/reference]

It imports the reference package from ``<reference>/src`` and records, for
small seeded inputs, the outputs of the reference functions the oracle
restates.  Every fixture is data: inputs and the reference's outputs, stored
as ``.npz`` (no pickles) next to this script.

Fixtures
  pairs.npz          cell_linked_pairs / brute_force_pairs / periodic_pairs
                     on seeded clouds (1-D..3-D, FP32 and FP64, duplicates
                     and at-cutoff points) - neighbor_search.py:29-120,
                     threed/neighbors3.py:71-170
  rhs2d.npz          droplet_impact.rhs on the small impact state with walls,
                     friction, Molteni/Antuono diffusion, split loads, FP64
                     and FP32 - droplet_impact.py:810-1228
  lattices.npz       make_particles / wall_lattice / make_column and the 3-D
                     lattices - droplet_impact.py:625-754, dam_break.py:143-169,
                     dam_break_3d.py:185-283, kleefsman_obstacle_3d.py:187-205
  run2d.npz          droplet_impact.run rows with output_every = 1 (meters per
                     step) and the final snapshot - droplet_impact.py:1607-1795
  threed_order4.npz  the same operators with the fourth-order viscous term
  kleefsman_gauges.npz  the Kleefsman gauge and tap operators on a cloud, and a
                     short run_case series; kleefsman_checkpoint_v3.pt a
                     checkpoint written by save_checkpoint
  kleefsman_adaptive.npz  Kleefsman run_case with adaptive_dt
  tvc.npz            transport_velocity_correction on a block over a wall and
                     dam_break.run_case with tvc_every (four variants, and one
                     with shifting and the adaptive step)
  lind_shift.npz     particle_shifting.shift_positions on a seeded column and a
                     dam_break.run_case with shift_every
  rhs2d_opts.npz     droplet_impact.rhs with its opt-in branches: wrap_width,
                     nu_field, density_offset, density renormalisation,
                     boundary-consistent density, surface-pressure closure
  threed.npz         threed.operators3.accelerations (periodic, renormalised,
                     walls + Adami, Antuono diffusion) - operators3.py:51-295
  rhs3d.npz          dam_break_3d.rhs_3d with wall forces - dam_break_3d.py:290-422
  deltaplus.npz      delta_plus_step with and without walls - delta_plus.py:105-229
  rhs2d_walls.npz    droplet_impact.rhs with wall_vel / wall_accel (FP64, FP32)
  rhs3d_opts.npz     dam_break_3d.rhs_3d: FP32, Molteni, closure, no walls, pairs=None
  dam3d_run.npz      dam_break_3d.run_case, fixed and adaptive step
"""
from __future__ import annotations

import math
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _np(t):
    return t.detach().cpu().numpy()


def main(reference: str = "/root/reference") -> None:
    sys.path.insert(0, str(Path(reference) / "src"))
    import torch
    torch.set_num_threads(1)
    from tensorsph.neighbor_search import (brute_force_pairs, cell_linked_pairs,
                                            PairListCache)
    from tensorsph.threed.neighbors3 import periodic_pairs
    from tensorsph import droplet_impact as di
    from tensorsph import dam_break as db
    from tensorsph import dam_break_3d as db3
    from tensorsph import kleefsman_obstacle_3d as kl
    from tensorsph.threed import operators3 as op3
    from tensorsph.threed import kernel3 as k3
    from tensorsph.threed.solver import cubic_lattice, wall_shell
    from tensorsph.delta_plus import delta_plus_step

    # ------------------------------------------------------------------ pairs
    out = {}
    case = 0
    for dim in (1, 2, 3):
        for seed in (0, 7, 42):
            for dtype in (torch.float64, torch.float32):
                g = torch.Generator().manual_seed(seed)
                pos = 4.0 * torch.randn((137, dim), generator=g, dtype=torch.float64)
                pos[1] = pos[0]
                pos[2] = pos[0]
                pos[2, 0] += 0.25
                pos = pos.to(dtype)
                out[f"c{case}_pos"] = _np(pos)
                out[f"c{case}_cutoff"] = np.array(0.25)
                out[f"c{case}_cell"] = _np(cell_linked_pairs(pos, 0.25))
                out[f"c{case}_brute"] = _np(brute_force_pairs(pos, 0.25))
                case += 1
    # lattice exactly at the cutoff (ties at d^2 == cutoff^2) in both dtypes
    for dtype in (torch.float64, torch.float32):
        ax = torch.arange(12, dtype=torch.float64) * 0.1
        pos = torch.cartesian_prod(ax, ax).to(dtype)
        out[f"c{case}_pos"] = _np(pos)
        out[f"c{case}_cutoff"] = np.array(0.1)
        out[f"c{case}_cell"] = _np(cell_linked_pairs(pos, 0.1))
        out[f"c{case}_brute"] = _np(brute_force_pairs(pos, 0.1))
        case += 1
    # uniform cloud, the reference's timing case shape (neighbor-search.md)
    g = torch.Generator().manual_seed(21)
    pos = torch.rand((1000, 2), generator=g, dtype=torch.float64)
    out[f"c{case}_pos"] = _np(pos)
    out[f"c{case}_cutoff"] = np.array(0.1)
    out[f"c{case}_cell"] = _np(cell_linked_pairs(pos, 0.1))
    out[f"c{case}_brute"] = _np(brute_force_pairs(pos, 0.1))
    case += 1
    out["ncases"] = np.array(case)
    # periodic (minimum image), jittered lattice
    n = 7
    dx = 1.0 / n
    lat = cubic_lattice((n, n, n), dx) + 0.5 * dx
    rng = np.random.Generator(np.random.PCG64(20261203))
    jit = torch.from_numpy(rng.uniform(-0.01 * dx, 0.01 * dx, size=lat.shape))
    lat = (lat + jit).remainder(1.0)
    cut = 2.0 * 1.3 * dx * 1.25
    out["per_pos"] = _np(lat)
    out["per_cutoff"] = np.array(cut)
    out["per_pairs"] = _np(periodic_pairs(lat, 1.0, cut))
    np.savez_compressed(HERE / "pairs.npz", **out)

    # ------------------------------------------------------------------ rhs2d
    out = {}
    cpu = torch.device("cpu")
    for tag, dtype in (("f64", torch.float64), ("f32", torch.float32)):
        torch.set_default_dtype(dtype)
        variants = {
            "base": {},
            "fric": {"wall_friction": 0.5},
            "molt": {"density_diffusion": 0.1},
            "antu": {"density_diffusion": 0.1, "density_diffusion_form": "antuono"},
        }
        for name, over in variants.items():
            cfg = di.Config(dx=0.05, pool_width=0.4, pool_depth=0.15, drop_radius=0.05,
                            drop_center_y=0.25, device="cpu", viscosity=0.01, **over)
            pos, mass = di.make_particles(cfg, cpu)
            wall = di.wall_lattice(cfg, cpu)
            g = torch.Generator().manual_seed(4)
            vel = torch.randn(pos.shape, generator=g, dtype=dtype) * 0.05
            # compress the lattice a little so pressure is non-zero everywhere
            pos = pos * 0.97
            h = cfg.h_factor * cfg.dx
            cache = PairListCache(2.0 * h, 0.25 * h, fluid_count=pos.shape[0])
            pairs, _ = cache.get(torch.cat((pos, wall)))
            rho, p, acc, loads = di.rhs(pos, vel, mass, cfg, pairs=pairs, wall=wall,
                                        split=True)
            key = f"{tag}_{name}"
            out[key + "_pos"] = _np(pos); out[key + "_vel"] = _np(vel)
            out[key + "_mass"] = _np(mass); out[key + "_wall"] = _np(wall)
            out[key + "_pairs"] = _np(pairs)
            out[key + "_rho"] = _np(rho); out[key + "_p"] = _np(p); out[key + "_acc"] = _np(acc)
            out[key + "_reaction"] = _np(loads["reaction"])
            out[key + "_face"] = _np(loads["face_pressure"])
            out[key + "_params"] = np.array([cfg.dx, cfg.h_factor, cfg.rho0, cfg.sound_speed,
                                             cfg.gamma, cfg.gravity, cfg.viscosity,
                                             cfg.wall_friction, cfg.dt, cfg.density_diffusion,
                                             1.0 if cfg.density_diffusion_form == "antuono" else 0.0,
                                             cfg.density_diffusion_eigenvalue_floor,
                                             float(di.lattice_mass_unit(cfg))])
            if name == "base":
                rho2, p2, acc2 = di.rhs(pos, vel, mass, cfg, wall=wall)   # pairs=None
                out[key + "_nopairs_rho"] = _np(rho2)
                out[key + "_nopairs_acc"] = _np(acc2)
        torch.set_default_dtype(torch.float64)
    # dense-reference case (no walls), test_pair_based_rhs.py:30-42
    g = torch.Generator().manual_seed(5)
    cfg = di.Config(dx=0.1, h_factor=1.3, rho0=1000.0, sound_speed=8.0, gamma=7.0,
                    gravity=9.81, viscosity=0.03, device="cpu")
    pos = torch.randn((61, 2), generator=g, dtype=torch.float64) * 0.12
    vel = torch.randn((61, 2), generator=g, dtype=torch.float64) * 0.2
    mass = (torch.rand((61,), generator=g, dtype=torch.float64) + 0.2) * 0.01
    rho, p, acc = di.rhs(pos, vel, mass, cfg)
    out["dense_pos"] = _np(pos); out["dense_vel"] = _np(vel); out["dense_mass"] = _np(mass)
    out["dense_rho"] = _np(rho); out["dense_p"] = _np(p); out["dense_acc"] = _np(acc)
    np.savez_compressed(HERE / "rhs2d.npz", **out)

    # --------------------------------------------------------------- lattices
    out = {}
    torch.set_default_dtype(torch.float64)
    cfg1 = di.Config(dx=0.005, pool_width=1.0, pool_depth=0.2, drop_radius=0.05,
                     drop_center_y=0.35, device="cpu", viscosity=1e-6,
                     sound_speed=10.0 * math.sqrt(2 * 9.81 * 0.2), dt=9.84e-5)
    pos, mass = di.make_particles(cfg1, cpu)
    out["impact_pos"] = _np(pos); out["impact_mass"] = _np(mass)
    out["impact_wall_top06"] = _np(di.wall_lattice(cfg1, cpu, top=0.6))
    out["impact_wall_default"] = _np(di.wall_lattice(cfg1, cpu))
    out["impact_mass_unit"] = np.array(di.lattice_mass_unit(cfg1))
    out["lattice_sum_13"] = np.array(di.lattice_kernel_sum(1.3 * 0.005, 0.005))
    cfg2 = db.DamBreakConfig(dx=0.05, tank_height=3.0, column_width=1.0, column_height=2.0,
                             pool_width=4.0, sound_speed=10.0 * math.sqrt(2 * 9.81 * 2.0))
    pos, mass = db.make_column(cfg2, cpu)
    out["dam_pos"] = _np(pos); out["dam_mass"] = _np(mass)
    out["dam_wall"] = _np(di.wall_lattice(cfg2, cpu, top=cfg2.tank_top()))
    cfg3 = db3.DamBreak3DConfig(dx=0.05, tank_height=0.6, tank_depth=0.3, column_width=0.15,
                                column_depth=0.15, column_height=0.3, pool_width=1.0)
    pos, mass = db3.make_column_3d(cfg3, cpu)
    out["dam3_pos"] = _np(pos); out["dam3_mass"] = _np(mass)
    out["dam3_wall"] = _np(db3.wall_lattice_3d(cfg3, cpu))
    out["k3_lattice_sum"] = np.array(k3.lattice_kernel_sum(0.1, 0.13))
    cfgk = kl.Kleefsman3DConfig(dx=0.04)
    out["kl_obstacle"] = _np(kl.obstacle_lattice_3d(cfgk, cpu))
    np.savez_compressed(HERE / "lattices.npz", **out)

    # ------------------------------------------------------------------ run2d
    out = {}
    torch.set_default_dtype(torch.float64)
    cfg = di.Config(dx=0.05, pool_width=0.4, pool_depth=0.15, drop_radius=0.05,
                    drop_center_y=0.25, device="cpu", viscosity=0.01, steps=40,
                    output_every=1, fail_fast=False, dt=9.0e-4)
    with tempfile.TemporaryDirectory() as tmp:
        rows = di.run(cfg, tmp, plot=False)
        snap = torch.load(Path(tmp) / f"particles_{cfg.steps:06d}.pt", weights_only=True)
    keys = ["step", "time", "mass_total", "momentum_x", "momentum_y", "rho_min", "rho_max",
            "rho_mean", "pressure_min", "pressure_max", "max_speed", "max_acceleration",
            "kinetic_energy", "y_min", "y_max"]
    out["row_keys"] = np.array(keys)
    out["rows"] = np.array([[float(r[k]) for k in keys] for r in rows])
    out["final_pos"] = _np(snap["position"]); out["final_rho"] = _np(snap["density"])
    out["params"] = np.array([cfg.dx, cfg.steps, cfg.dt, cfg.viscosity])
    np.savez_compressed(HERE / "run2d.npz", **out)

    # ----------------------------------------------------------------- threed
    out = {}
    n = 8
    dx = 1.0 / n
    h = 1.3 * dx
    lat = cubic_lattice((n, n, n), dx) + 0.5 * dx
    rng = np.random.Generator(np.random.PCG64(20261203))
    lat = (lat + torch.from_numpy(rng.uniform(-0.01 * dx, 0.01 * dx, size=lat.shape))).remainder(1.0)
    mass = torch.full((lat.shape[0],), k3.calibrated_particle_mass(dx, h, 1.0))
    vel = torch.zeros_like(lat)
    vel[:, 0] = 1e-3 * torch.sin(2 * math.pi * lat[:, 2])
    vel[:, 1] = 3e-4 * torch.cos(2 * math.pi * lat[:, 0])
    pairs = periodic_pairs(lat, 1.0, 2.0 * h * 1.25)
    acc, rho, p = op3.accelerations(lat, vel, mass, h, pairs, rho0=1.0, sound_speed=10.0,
                                    gamma=7.0, viscosity=0.01, box=1.0, renormalise=True)
    out.update(per_pos=_np(lat), per_vel=_np(vel), per_mass=_np(mass), per_pairs=_np(pairs),
               per_acc=_np(acc), per_rho=_np(rho), per_p=_np(p), per_h=np.array(h))
    acc, rho, p = op3.accelerations(lat, vel, mass, h, pairs, rho0=1.0, sound_speed=10.0,
                                    gamma=7.0, viscosity=0.01, box=1.0, renormalise=False)
    out.update(per_raw_acc=_np(acc))
    # walls + Adami + gravity + Antuono diffusion (hydrostatic-like block)
    nf = (5, 5, 4)
    dxw = 0.1
    hw = 1.3 * dxw
    fluid = cubic_lattice(nf, dxw, origin=(3 * dxw, 3 * dxw, 3 * dxw))
    shell = wall_shell(nf, 3, dxw)
    g2 = torch.Generator().manual_seed(9)
    fluid = fluid + 0.02 * dxw * torch.randn(fluid.shape, generator=g2)
    allp = torch.cat((fluid, shell))
    m = torch.full((allp.shape[0],), k3.calibrated_particle_mass(dxw, hw, 1000.0))
    m[: fluid.shape[0]] *= 1.0 + 0.01 * torch.rand(fluid.shape[0], generator=g2)
    v = torch.zeros_like(allp)
    v[: fluid.shape[0]] = 0.05 * torch.randn(fluid.shape, generator=g2)
    pw = periodic_pairs  # noqa: F841  (walls use the open search)
    pairs = cell_linked_pairs(allp, 2.0 * hw)
    grav = [0.0, 0.0, -9.81]
    acc, rho, p = op3.accelerations(allp, v, m, hw, pairs, rho0=1000.0, sound_speed=20.0,
                                    gamma=7.0, viscosity=1e-3, gravity=grav,
                                    n_fluid=fluid.shape[0], renormalise=True,
                                    density_diffusion=0.1, dt=1e-4,
                                    density_diffusion_form="antuono")
    out.update(wall_pos=_np(allp), wall_vel=_np(v), wall_mass=_np(m), wall_pairs=_np(pairs),
               wall_nf=np.array(fluid.shape[0]), wall_h=np.array(hw),
               wall_acc=_np(acc), wall_rho=_np(rho), wall_p=_np(p))
    np.savez_compressed(HERE / "threed.npz", **out)

    # ------------------------------------------------------------------ rhs3d
    out = {}
    cfg3 = db3.DamBreak3DConfig(dx=0.05, tank_height=0.4, tank_depth=0.2, column_width=0.2,
                                column_depth=0.2, column_height=0.25, pool_width=0.4,
                                viscosity=1e-3, wall_friction=0.3)
    pos, mass = db3.make_column_3d(cfg3, cpu)
    wall = db3.wall_lattice_3d(cfg3, cpu)
    g3 = torch.Generator().manual_seed(13)
    vel = 0.1 * torch.randn(pos.shape, generator=g3)
    pos = pos + 0.01 * cfg3.dx * torch.randn(pos.shape, generator=g3)
    h = cfg3.h_factor * cfg3.dx
    cache = PairListCache(2.0 * h, 0.25 * h, fluid_count=pos.shape[0])
    pairs, _ = cache.get(torch.cat((pos, wall)))
    rho, p, acc, reaction = db3.rhs_3d(pos, vel, mass, cfg3, pairs=pairs, wall=wall,
                                       return_wall_force=True)
    out.update(pos=_np(pos), vel=_np(vel), mass=_np(mass), wall=_np(wall), pairs=_np(pairs),
               rho=_np(rho), p=_np(p), acc=_np(acc), reaction=_np(reaction),
               params=np.array([cfg3.dx, cfg3.h_factor, cfg3.rho0, cfg3.sound_speed,
                                cfg3.gamma, cfg3.gravity, cfg3.viscosity, cfg3.wall_friction,
                                cfg3.dt, db3._mass_3d(cfg3.dx, cfg3.h_factor, cfg3.rho0, True)]))
    np.savez_compressed(HERE / "rhs3d.npz", **out)

    # -------------------------------------------------------------- deltaplus
    out = {}
    DX = 0.01
    axis = (torch.arange(15, dtype=torch.float64) + 0.5) * DX
    pos = torch.cartesian_prod(axis, axis)
    pos[7 * 15 + 7, 0] += 0.23 * DX
    pos[7 * 15 + 13, 0] -= 0.2 * DX
    pos[99] += 0.2 * DX
    g4 = torch.Generator().manual_seed(17)
    vel = 0.05 * torch.randn(pos.shape, generator=g4, dtype=torch.float64)
    rho = 1000.0 + 0.5 * torch.randn(pos.shape[0], generator=g4, dtype=torch.float64)
    mass = rho * DX ** 2
    # floor and left wall ghost stations with normals (spheric19.py:124-131 style)
    xs = (torch.arange(-3, 16, dtype=torch.float64) + 0.5) * DX
    floor = torch.stack((xs, torch.full_like(xs, -0.5 * DX)), 1)
    ys = (torch.arange(0, 16, dtype=torch.float64) + 0.5) * DX
    left = torch.stack((torch.full_like(ys, -0.5 * DX), ys), 1)
    wall = torch.cat((floor, left))
    wn = torch.cat((torch.tensor([[0.0, 1.0]]).expand(len(floor), 2),
                    torch.tensor([[1.0, 0.0]]).expand(len(left), 2)))
    wv = torch.zeros_like(wall)
    wv[len(floor):, 1] = 0.1
    params = dict(dx=DX, sound_speed=20.0, reference_speed=0.5, dt=1e-4, rho0=1000.0,
                  gravity=9.81, viscosity=1e-6, diffusion=0.1, alpha=0.02)
    pairs = cell_linked_pairs(torch.cat((pos, wall)), 4 * DX * 1.1)
    step = delta_plus_step(pos, vel, rho, mass, wall=wall, wall_velocity=wv,
                           wall_normal=wn, pairs=pairs, **params)
    out.update(pos=_np(pos), vel=_np(vel), rho=_np(rho), mass=_np(mass), wall=_np(wall),
               wall_vel=_np(wv), wall_normal=_np(wn), pairs=_np(pairs),
               params=np.array([params[k] for k in ("dx", "sound_speed", "reference_speed",
                                                    "dt", "rho0", "gravity", "viscosity",
                                                    "diffusion", "alpha")]),
               out_pos=_np(step.position), out_vel=_np(step.velocity),
               out_rho=_np(step.density), out_acc=_np(step.acceleration),
               out_reaction=_np(step.wall_reaction), out_shift=_np(step.shift))
    # free patch (no walls): the elliptical-drop initial state at 8 cells/radius
    cells = 8
    dxe = 1 / cells
    ax = (torch.arange(-cells, cells, dtype=torch.float64) + 0.5) * dxe
    pe = torch.cartesian_prod(ax, ax)
    r2 = pe.square().sum(1)
    pe = pe[r2 < 1]
    de = 1000 * (1 + 0.5 * (1 - r2[r2 < 1]) / 20.0 ** 2)
    me = de * dxe ** 2
    ve = pe * pe.new_tensor([1.0, -1.0])
    pairs_e = cell_linked_pairs(pe, 4 * dxe + 0.4 * dxe)
    st = delta_plus_step(pe, ve, de, me, dx=dxe, dt=1e-3, sound_speed=20.0,
                         reference_speed=1.0, gravity=0.0, viscosity=0.0, pairs=pairs_e)
    out.update(free_pos=_np(pe), free_vel=_np(ve), free_rho=_np(de), free_mass=_np(me),
               free_pairs=_np(pairs_e), free_out_pos=_np(st.position),
               free_out_vel=_np(st.velocity), free_out_rho=_np(st.density),
               free_out_shift=_np(st.shift))
    np.savez_compressed(HERE / "deltaplus.npz", **out)
    print("golden fixtures written to", HERE)


def shear(reference: str = "/root/reference") -> None:
    """run_shear_wave (threed/solver.py:211-308) at n = 10 with an amplitude
    large enough that the skin test rebuilds the pair list inside the run."""
    sys.path.insert(0, str(Path(reference) / "src"))
    from tensorsph.threed.solver import run_shear_wave
    res = run_shear_wave(n=10, amplitude=0.5, steps=60, sample_every=5)
    keys = ("mass", "kinetic_energy", "internal_energy", "total_energy", "momentum_magnitude")
    rows = np.array([[d[k] for k in keys] + list(d["momentum"]) + list(d["angular_momentum"])
                     + list(d["centre_of_mass"]) for d in res.diagnostics])
    np.savez_compressed(HERE / "shear.npz", times=np.array(res.times),
                        amplitudes=np.array(res.amplitudes), rows=rows,
                        row_keys=np.array(list(keys) + ["px", "py", "pz", "lx", "ly", "lz",
                                                        "cx", "cy", "cz"]),
                        pairs_last=np.int64(res.pairs_last),
                        pair_rebuilds=np.int64(res.pair_rebuilds),
                        momentum_scale=np.float64(res.momentum_scale),
                        fit_rate=np.float64(res.fit_rate),
                        params=np.array([res.n, res.dx, res.h, res.nu, res.wavenumber,
                                         res.sound_speed, res.amplitude0, res.dt, res.steps]))
    print("shear fixture written to", HERE)


def order4(reference: str = "/root/reference") -> None:
    """threed.operators3 with the fourth-order viscous correction
    (operators3.py:146-151): accelerations(viscous_order=4) and
    viscous_acceleration(order=4) on a jittered periodic lattice with
    unequal masses, renormalised and not, and a short shear-wave run with
    viscous_order=4."""
    sys.path.insert(0, str(Path(reference) / "src"))
    import torch
    from tensorsph.threed.neighbors3 import periodic_pairs
    from tensorsph.threed import operators3 as op3
    from tensorsph.threed import kernel3 as k3
    from tensorsph.threed.solver import cubic_lattice, run_shear_wave
    torch.set_default_dtype(torch.float64)
    n = 8
    dx = 1.0 / n
    h = 1.3 * dx
    lat = cubic_lattice((n, n, n), dx) + 0.5 * dx
    rng = np.random.Generator(np.random.PCG64(20261204))
    lat = (lat + torch.from_numpy(rng.uniform(-0.05 * dx, 0.05 * dx, size=lat.shape))).remainder(1.0)
    mass = torch.full((lat.shape[0],), k3.calibrated_particle_mass(dx, h, 1.0))
    mass = mass * (1.0 + 0.01 * torch.from_numpy(rng.uniform(-1.0, 1.0, size=mass.shape)))
    vel = torch.zeros_like(lat)
    vel[:, 0] = 1e-3 * torch.sin(2 * math.pi * lat[:, 2])
    vel[:, 1] = 3e-4 * torch.cos(2 * math.pi * lat[:, 0])
    vel[:, 2] = 2e-4 * torch.sin(4 * math.pi * lat[:, 1])
    pairs = periodic_pairs(lat, 1.0, 2.0 * h * 1.25)
    out = dict(pos=_np(lat), vel=_np(vel), mass=_np(mass), pairs=_np(pairs), h=np.array(h))
    for tag, ren in (("ren", True), ("raw", False)):
        acc, rho, p = op3.accelerations(lat, vel, mass, h, pairs, rho0=1.0, sound_speed=10.0,
                                        gamma=7.0, viscosity=0.01, box=1.0, renormalise=ren,
                                        viscous_order=4)
        out[f"{tag}_acc"] = _np(acc)
        out[f"{tag}_rho"] = _np(rho)
        out[f"{tag}_visc"] = _np(op3.viscous_acceleration(lat, vel, mass, rho, 0.01, h, pairs, box=1.0,
                                                          renormalise=ren, order=4))
    # run_shear_wave with the fourth-order operator (solver.py:211-308)
    res = run_shear_wave(n=10, amplitude=0.5, steps=60, sample_every=5, viscous_order=4)
    out["shear_times"] = np.array(res.times)
    out["shear_amplitudes"] = np.array(res.amplitudes)
    out["shear_ke"] = np.array([d["kinetic_energy"] for d in res.diagnostics])
    out["shear_momentum"] = np.array([list(d["momentum"]) for d in res.diagnostics])
    out["shear_rebuilds"] = np.int64(res.pair_rebuilds)
    np.savez_compressed(HERE / "threed_order4.npz", **out)
    print("order-4 fixture written to", HERE)


def rhs_options(reference: str = "/root/reference") -> None:
    """droplet_impact.rhs with its opt-in branches (droplet_impact.py:810-1228):
    wrap_width (x-periodic strip, :921-927), nu_field (:1164-1178),
    density_offset (:936-943), density_renormalization (:954-964),
    boundary_consistent_density (:1090-1112) and the surface-pressure closure
    (:1031-1038, :429-492, both donor sets), on the small impact state."""
    sys.path.insert(0, str(Path(reference) / "src"))
    import torch
    from tensorsph import droplet_impact as di
    from tensorsph.neighbor_search import PairListCache
    torch.set_default_dtype(torch.float64)
    cpu = torch.device("cpu")
    out = {}
    base = dict(dx=0.05, pool_width=0.4, pool_depth=0.15, drop_radius=0.05, drop_center_y=0.25,
                device="cpu", viscosity=0.01)
    variants = {
        "offset": ({}, "offset"),
        "shepard": ({"density_renormalization": True}, None),
        "nufield": ({"wall_friction": 0.3}, "nu"),
        "bcd": ({"boundary_consistent_density": True}, None),
        "closure": ({"surface_pressure_closure": True}, None),
        "closure_all": ({"surface_pressure_closure": True, "surface_pressure_donors": "all"}, None),
        "closure3": ({"surface_pressure_closure": True, "surface_pressure_layers": 3}, None),
        "closure_bcd": ({"surface_pressure_closure": True, "boundary_consistent_density": True,
                         "density_diffusion": 0.1, "dt": 1e-4}, None),
        "shepard_diff": ({"density_renormalization": True, "density_diffusion": 0.1,
                          "density_diffusion_form": "molteni", "dt": 1e-4}, None),
    }
    for name, (over, extra) in variants.items():
        cfg = di.Config(**base, **over)
        pos, mass = di.make_particles(cfg, cpu)
        wall = di.wall_lattice(cfg, cpu)
        g = torch.Generator().manual_seed(11)
        vel = torch.randn(pos.shape, generator=g) * 0.05
        pos = pos + 0.05 * cfg.dx * torch.randn(pos.shape, generator=g)
        if name.startswith("shepard"):
            # with equal masses the Shepard-normalised density is rho0 to round-off
            mass = mass * (1.0 + 0.05 * torch.rand(mass.shape, generator=g))
        h = cfg.h_factor * cfg.dx
        cache = PairListCache(2.0 * h, 0.25 * h, fluid_count=pos.shape[0])
        pairs, _ = cache.get(torch.cat((pos, wall)))
        kw = {}
        if extra == "offset":
            kw["density_offset"] = 5.0 * torch.randn(pos.shape[0], generator=g)
            out[name + "_offset"] = _np(kw["density_offset"])
        if extra == "nu":
            kw["nu_field"] = 0.005 + 0.015 * torch.rand(pos.shape[0], generator=g)
            out[name + "_nu"] = _np(kw["nu_field"])
        rho, p, acc, loads = di.rhs(pos, vel, mass, cfg, pairs=pairs, wall=wall, split=True,
                                    dt=cfg.dt, **kw)
        out.update({name + "_pos": _np(pos), name + "_vel": _np(vel), name + "_mass": _np(mass),
                    name + "_wall": _np(wall), name + "_pairs": _np(pairs), name + "_rho": _np(rho),
                    name + "_p": _np(p), name + "_acc": _np(acc),
                    name + "_reaction": _np(loads["reaction"]),
                    name + "_face": _np(loads["face_pressure"])})
        out[name + "_lattice_sum"] = np.array(di.lattice_kernel_sum(h, cfg.dx))
        out[name + "_unit"] = np.array(float(di.lattice_mass_unit(cfg)))
    # x-periodic strip: bottom wall rows only, pairs by a minimum-image brute force
    cfg = di.Config(**base)
    pos, mass = di.make_particles(cfg, cpu)
    wall = di.wall_lattice(cfg, cpu)
    width = cfg.pool_width
    keep = (wall[:, 1] < 0.0) & (wall[:, 0] >= 0.0) & (wall[:, 0] < width)
    wall = wall[keep]
    g = torch.Generator().manual_seed(12)
    pos = pos + 0.05 * cfg.dx * torch.randn(pos.shape, generator=g)
    pos[:, 0] = pos[:, 0].remainder(width)
    vel = torch.randn(pos.shape, generator=g) * 0.05
    h = cfg.h_factor * cfg.dx
    allp = _np(torch.cat((pos, wall)))
    n = pos.shape[0]
    d = allp[:, None, :] - allp[None, :, :]
    d[..., 0] -= width * np.round(d[..., 0] / width)
    close = (d ** 2).sum(-1) <= (2.0 * h * 1.1) ** 2
    ii, jj = np.nonzero(np.triu(close, 1))
    sel = (ii < n) | (jj < n)
    pairs = torch.from_numpy(np.stack((ii[sel], jj[sel]), 1).astype(np.int64))
    rho, p, acc, loads = di.rhs(pos, vel, mass, cfg, pairs=pairs, wall=wall, split=True,
                                wrap_width=width)
    out.update({"wrap_pos": _np(pos), "wrap_vel": _np(vel), "wrap_mass": _np(mass),
                "wrap_wall": _np(wall), "wrap_pairs": _np(pairs), "wrap_rho": _np(rho),
                "wrap_p": _np(p), "wrap_acc": _np(acc), "wrap_reaction": _np(loads["reaction"]),
                "wrap_face": _np(loads["face_pressure"]), "wrap_width": np.array(width),
                "wrap_unit": np.array(float(di.lattice_mass_unit(cfg)))})
    np.savez_compressed(HERE / "rhs2d_opts.npz", **out)
    print("rhs option fixture written to", HERE)


def kleefsman_gauges(reference: str = "/root/reference") -> None:
    """The gauge operators of kleefsman_obstacle_3d / marin_probes on a seeded
    particle cloud around the box (slab_height, peak_height, bypass_count
    :273-318, centreline_height and tap_node_pressures with point and disk
    taps marin_probes.py:25-79), a run_case at dx = 0.05 to t = 0.62 s (past the impact on the box) with its
    on_sample series (:553-726), and a v3 checkpoint written by
    save_checkpoint (:462-517)."""
    sys.path.insert(0, str(Path(reference) / "src"))
    import torch
    from tensorsph import kleefsman_obstacle_3d as kl
    from tensorsph import marin_probes as mp
    torch.set_default_dtype(torch.float64)
    out = {}
    cfg = kl.Kleefsman3DConfig(dx=0.05, t_end=0.62, output_every=25, pressure_disk_diameter=0.03)
    g = torch.Generator().manual_seed(21)
    n = 6000
    lo = torch.tensor([2.0, 0.0, 0.0])
    hi = torch.tensor([2.7, 0.35, 1.0])
    pos = lo + (hi - lo) * torch.rand((n, 3), generator=g)
    inside = ((pos[:, 0] > cfg.obstacle_x) & (pos[:, 0] < cfg.obstacle_x1())
              & (pos[:, 1] < cfg.obstacle_height)
              & (pos[:, 2] > cfg.obstacle_z) & (pos[:, 2] < cfg.obstacle_z1()))
    pos = pos[~inside]
    m = pos.shape[0]
    rho = 1000.0 * (1.0 + 0.02 * torch.rand(m, generator=g))
    pressure = 2000.0 * torch.rand(m, generator=g)
    volume = cfg.dx ** 3 * (1.0 + 0.1 * torch.rand(m, generator=g))
    out.update(cloud_pos=pos.numpy(), cloud_rho=rho.numpy(), cloud_p=pressure.numpy(),
               cloud_vol=volume.numpy())
    out["cloud_slab"] = np.array(kl.slab_height(pos, cfg))
    out["cloud_peak"] = np.array(kl.peak_height(pos, cfg))
    out["cloud_bypass"] = np.array(kl.bypass_count(pos, cfg))
    out["cloud_h2"] = np.array(float(mp.centreline_height(pos, volume, cfg)))
    taps = mp.sensor_positions(cfg, pos)
    for tag, dia in (("point", 0.0), ("disk", 0.03)):
        nodes = mp.tap_disk_nodes(cfg, taps, dia)
        p, support = mp.tap_node_pressures(pos, pressure, rho, volume, cfg, nodes)
        out[f"cloud_{tag}_nodes"] = nodes.numpy()
        out[f"cloud_{tag}_p"] = p.numpy()
        out[f"cloud_{tag}_support"] = support.numpy()
    out["cloud_disk_avg"] = mp.disk_average(torch.from_numpy(out["cloud_disk_p"])).numpy()
    # a short run with the per-sample series
    samples = []
    run = kl.run_case(cfg, on_sample=lambda i, s: samples.append(s))
    keys = ["t_s", "slab_height_m", "height_operator_m", "front_x_m", "n_bypass", "H2_local_m"]
    out["run_rows"] = np.array([[float(s[k]) for k in keys] for s in samples])
    out["run_row_keys"] = np.array(keys)
    out["run_force"] = np.array([s["obstacle_sph_force_N"] for s in samples])
    out["run_impulse"] = np.array([s["obstacle_sph_impulse_Ns"] for s in samples])
    out["run_moment"] = np.array([s["obstacle_sph_moment_Nm"] for s in samples])
    out["run_tap_p"] = np.array([s["tap_pressure_Pa"] for s in samples])
    out["run_tap_support"] = np.array([s["tap_support"] for s in samples])
    out["run_disk_p"] = np.array([s["disk_pressure_Pa"] for s in samples])
    summary = run.to_dict()
    for k in ("peak_height_m", "peak_time_s", "peak_height_operator_m", "max_speed",
              "max_density_compression", "steps", "simulated_time_s", "local_h2_peak_m",
              "boundary_escapes", "obstacle_clamps", "max_bypass_particles"):
        v = summary[k]
        out["summary_" + k] = np.array(np.nan if v is None else float(v))
    for k in ("arrival_time_s", "local_h2_arrival_s"):
        v = summary[k]
        out["summary_" + k] = np.array(np.nan if v is None else float(v))
    out["summary_sph_impulse"] = np.array(summary["obstacle_sph_impulse_Ns"])
    out["summary_clamp_impulse"] = np.array(summary["obstacle_clamp_impulse_Ns"])
    out["params"] = np.array([cfg.dx, cfg.t_end, cfg.dt, cfg.output_every,
                              cfg.pressure_disk_diameter, cfg.sound_speed])
    np.savez_compressed(HERE / "kleefsman_gauges.npz", **out)
    # a checkpoint file as save_checkpoint writes it (tensors, lists, numbers)
    ck = kl.Kleefsman3DRun(cfg=cfg)
    ck.times, ck.slab, ck.height = [0.0, 0.1], [0.0, 0.05], [0.0, 0.12]
    ck.local_h2, ck.front, ck.bypass = [0.0, 0.04], [1.2, 2.0], [0.0, 3.0]
    kl.save_checkpoint(HERE / "kleefsman_checkpoint_v3.pt", cfg, step=40, elapsed=40 * cfg.dt,
                       dt=cfg.dt, positions=pos[:50], velocity=0.1 * pos[:50], run=ck,
                       speeds=1.5, compression=0.01, escapes=0, clamps=2)
    print("kleefsman fixtures written to", HERE)


def lind_shift(reference: str = "/root/reference") -> None:
    """particle_shifting.shift_positions (particle_shifting.py:166-224) on a
    seeded 2-D column with a free surface, and dam_break.run_case with
    shift_every (dam_break.py:424-433)."""
    sys.path.insert(0, str(Path(reference) / "src"))
    import torch
    from tensorsph import dam_break as db
    from tensorsph.particle_shifting import shift_positions
    torch.set_default_dtype(torch.float64)
    out = {}
    g = torch.Generator().manual_seed(31)
    dx = 0.02
    h = 1.3 * dx
    xs = (torch.arange(20) + 0.5) * dx
    ys = (torch.arange(12) + 0.5) * dx
    pos = torch.stack(torch.meshgrid(xs, ys, indexing="ij"), -1).reshape(-1, 2)
    pos = pos + 0.08 * dx * torch.randn(pos.shape, generator=g)
    volume = dx * dx * (1.0 + 0.05 * torch.rand(pos.shape[0], generator=g))
    for tag, kw in (("a", dict(dt=1e-3, v_ref=2.0)), ("clip", dict(dt=1e-2, v_ref=5.0, limit=0.05)),
                    ("strong", dict(dt=2e-3, v_ref=1.0, strength=3.0, free_surface=0.9))):
        new, st = shift_positions(pos, volume, h, return_state=True, **kw)
        out[f"{tag}_out"] = new.numpy()
        out[f"{tag}_c"] = st.concentration.numpy()
        out[f"{tag}_cref"] = np.array(st.c_reference)
        out[f"{tag}_fs"] = st.free_surface.numpy()
        out[f"{tag}_kw"] = np.array([kw["dt"], kw["v_ref"], kw.get("strength", 1.0), kw.get("limit", 0.25),
                                     kw.get("free_surface", 0.95)])
    out.update(pos=pos.numpy(), volume=volume.numpy(), h=np.array(h))
    cfg = db.DamBreakConfig(dx=0.02, tank_height=0.6, column_width=0.3, column_height=0.4,
                            pool_width=0.5, sound_speed=20.0, dt=3.0e-4, viscosity=1e-3, steps=200,
                            output_every=20, shift_every=5, device="cpu")
    res = db.run_case(cfg)
    out["run_times"] = np.array(res.times)
    out["run_front"] = np.array(res.front)
    out["run_ke"] = np.array(res.kinetic_energy)
    out["run_pe"] = np.array(res.potential_energy)
    out["run_final"] = res.final_positions.numpy()
    out["run_rebuilds"] = np.array(res.rebuilds)
    np.savez_compressed(HERE / "lind_shift.npz", **out)
    print("shifting fixture written to", HERE)


def tvc(reference: str = "/root/reference") -> None:
    """transport_velocity_correction (transport_velocity_correction.py:285-513)
    on a seeded fluid block with a wall layer, and dam_break.run_case with
    tvc_every (dam_break.py:434-460) in three variants."""
    sys.path.insert(0, str(Path(reference) / "src"))
    import torch
    from tensorsph import dam_break as db
    from tensorsph.neighbor_search import cell_linked_pairs
    from tensorsph.transport_velocity_correction import (transport_velocity_correction,
                                                         truncated_linear)
    torch.set_default_dtype(torch.float64)
    out = {}
    g = torch.Generator().manual_seed(41)
    dx = 0.02
    h = 1.3 * dx
    xs = (torch.arange(16) + 0.5) * dx
    ys = (torch.arange(10) + 0.5) * dx
    fluid = torch.stack(torch.meshgrid(xs, ys, indexing="ij"), -1).reshape(-1, 2)
    fluid = fluid + 0.1 * dx * torch.randn(fluid.shape, generator=g)
    wx = (torch.arange(-3, 19) + 0.5) * dx
    wall = torch.stack(torch.meshgrid(wx, torch.tensor([-0.5 * dx, -1.5 * dx]), indexing="ij"), -1).reshape(-1, 2)
    allp = torch.cat((fluid, wall))
    vol = torch.cat((dx * dx * (1.0 + 0.05 * torch.rand(fluid.shape[0], generator=g)),
                     torch.full((wall.shape[0],), dx * dx)))
    pairs = cell_linked_pairs(allp, 2.0 * h)
    for tag, kw in (("plain", {}), ("lim", dict(alpha=0.1, limiter=truncated_linear(100.0), max_shift=0.002)),
                    ("scope", dict(alpha=0.05, scope_tol=0.5))):
        r = transport_velocity_correction(allp, vol, h, pairs=pairs, wall_count=wall.shape[0], **kw)
        out[f"{tag}_g"] = r.gradient_integral.numpy()
        out[f"{tag}_shift"] = r.shift.numpy()
    out.update(pos=allp.numpy(), vol=vol.numpy(), nf=np.array(fluid.shape[0]), h=np.array(h),
               pairs=pairs.numpy())
    base = dict(dx=0.02, tank_height=0.6, column_width=0.3, column_height=0.4, pool_width=0.5,
                sound_speed=20.0, dt=3.0e-4, viscosity=1e-3, steps=200, output_every=20, device="cpu")
    for tag, kw in (("run_all", dict(tvc_every=5)),
                    ("run_lim", dict(tvc_every=4, tvc_limiter="truncated_linear", tvc_max_shift=0.01,
                                     tvc_alpha=0.1)),
                    ("run_bulkg", dict(tvc_every=5, tvc_scope="bulk_g", tvc_bulk_g=0.05)),
                    ("run_bulk", dict(tvc_every=5, tvc_scope="bulk", shift_every=10)),
                    ("run_adapt", dict(tvc_every=5, shift_every=7, adaptive_dt=True, max_courant=0.3))):
        res = db.run_case(db.DamBreakConfig(**base, **kw))
        out[f"{tag}_front"] = np.array(res.front)
        out[f"{tag}_ke"] = np.array(res.kinetic_energy)
        out[f"{tag}_final"] = res.final_positions.numpy()
        out[f"{tag}_tvc"] = np.array([res.max_tvc_shift, res.max_tvc_clipped, res.escaped, res.rebuilds])
        out[f"{tag}_times"] = np.array(res.times)
        out[f"{tag}_steps"] = np.array(res.steps_taken)
    np.savez_compressed(HERE / "tvc.npz", **out)
    print("tvc fixture written to", HERE)


def kleefsman_adaptive(reference: str = "/root/reference") -> None:
    """kleefsman_obstacle_3d.run_case with adaptive_dt (:644-713) at dx = 0.05,
    a Courant number that lets the step shrink with the flow speed."""
    sys.path.insert(0, str(Path(reference) / "src"))
    import torch
    from tensorsph import kleefsman_obstacle_3d as kl
    torch.set_default_dtype(torch.float64)
    cfg = kl.Kleefsman3DConfig(dx=0.05, t_end=0.45, output_every=25, adaptive_dt=True,
                               max_courant=0.085, max_steps=4000)
    samples = []
    run = kl.run_case(cfg, on_sample=lambda i, s: samples.append(s))
    d = run.to_dict()
    keys = ["t_s", "slab_height_m", "height_operator_m", "front_x_m", "n_bypass", "H2_local_m"]
    out = {"rows": np.array([[float(s[k]) for k in keys] for s in samples]), "row_keys": np.array(keys),
           "force": np.array([s["obstacle_sph_force_N"] for s in samples]),
           "impulse": np.array([s["obstacle_sph_impulse_Ns"] for s in samples]),
           "steps": np.array(d["steps"]), "simulated_time": np.array(d["simulated_time_s"]),
           "sph_impulse": np.array(d["obstacle_sph_impulse_Ns"]), "max_speed": np.array(d["max_speed"]),
           "compression": np.array(d["max_density_compression"]),
           "clamps": np.array(d["obstacle_clamps"]),
           "clamp_impulse": np.array(d["obstacle_clamp_impulse_Ns"]),
           "escapes": np.array(d["boundary_escapes"])}
    np.savez_compressed(HERE / "kleefsman_adaptive.npz", **out)
    print("adaptive kleefsman fixture written to", HERE)


def moving_walls(reference: str = "/root/reference") -> None:
    """droplet_impact.rhs with moving walls (wall_vel :1119-1131 enters the
    viscous and friction terms; wall_accel :1063-1072 enters the Adami body
    term, which the reference adds only when gravity != 0), FP64 and FP32,
    with split loads; and dam_break_3d.rhs_3d variants (FP32, Molteni
    diffusion, the surface-pressure closure, no walls, pairs=None)."""
    sys.path.insert(0, str(Path(reference) / "src"))
    import torch
    from tensorsph import droplet_impact as di
    from tensorsph import dam_break_3d as db3
    from tensorsph.neighbor_search import PairListCache
    cpu = torch.device("cpu")
    out = {}
    base = dict(dx=0.05, pool_width=0.4, pool_depth=0.15, drop_radius=0.05, drop_center_y=0.25,
                device="cpu", viscosity=0.01)
    variants = {
        "vel": ({"wall_friction": 0.4}, True, False),
        "acc": ({}, False, True),
        "both": ({"wall_friction": 0.25}, True, True),
        "acc_nograv": ({"gravity": 0.0}, False, True),
    }
    for tag, dtype in (("f64", torch.float64), ("f32", torch.float32)):
        torch.set_default_dtype(dtype)
        for name, (over, with_vel, with_acc) in variants.items():
            cfg = di.Config(**base, **over)
            pos, mass = di.make_particles(cfg, cpu)
            wall = di.wall_lattice(cfg, cpu)
            g = torch.Generator().manual_seed(21)
            vel = torch.randn(pos.shape, generator=g, dtype=dtype) * 0.05
            pos = pos * 0.97
            nw = wall.shape[0]
            kw = {}
            if with_vel:
                kw["wall_vel"] = torch.randn((nw, 2), generator=g, dtype=dtype) * 0.3
            if with_acc:
                kw["wall_accel"] = torch.randn((nw, 2), generator=g, dtype=dtype) * 4.0
            h = cfg.h_factor * cfg.dx
            cache = PairListCache(2.0 * h, 0.25 * h, fluid_count=pos.shape[0])
            pairs, _ = cache.get(torch.cat((pos, wall)))
            rho, p, acc, loads = di.rhs(pos, vel, mass, cfg, pairs=pairs, wall=wall, split=True,
                                        **kw)
            key = f"{tag}_{name}"
            out.update({key + "_pos": _np(pos), key + "_vel": _np(vel), key + "_mass": _np(mass),
                        key + "_wall": _np(wall), key + "_pairs": _np(pairs),
                        key + "_rho": _np(rho), key + "_p": _np(p), key + "_acc": _np(acc),
                        key + "_reaction": _np(loads["reaction"]),
                        key + "_face": _np(loads["face_pressure"])})
            for k, v in kw.items():
                out[key + "_" + k] = _np(v)
            out[key + "_params"] = np.array([cfg.dx, cfg.h_factor, cfg.rho0, cfg.sound_speed,
                                             cfg.gamma, cfg.gravity, cfg.viscosity,
                                             cfg.wall_friction, cfg.dt,
                                             float(di.lattice_mass_unit(cfg))])
    torch.set_default_dtype(torch.float64)
    np.savez_compressed(HERE / "rhs2d_walls.npz", **out)

    out = {}
    geo = dict(dx=0.05, tank_height=0.4, tank_depth=0.2, column_width=0.2, column_depth=0.2,
               column_height=0.25, pool_width=0.4, viscosity=1e-3)
    variants3 = {
        "f32": ({"wall_friction": 0.3}, torch.float32, True, True),
        "molt": ({"density_diffusion": 0.1, "dt": 1e-4}, torch.float64, True, True),
        "closure": ({"surface_pressure_closure": True, "wall_friction": 0.2}, torch.float64,
                    True, True),
        "nowall": ({}, torch.float64, False, True),
        "nopairs": ({"wall_friction": 0.3}, torch.float64, True, False),
    }
    for name, (over, dtype, with_wall, with_pairs) in variants3.items():
        torch.set_default_dtype(dtype)
        cfg3 = db3.DamBreak3DConfig(**geo, **over)
        pos, mass = db3.make_column_3d(cfg3, cpu)
        wall = db3.wall_lattice_3d(cfg3, cpu) if with_wall else None
        g3 = torch.Generator().manual_seed(31)
        vel = 0.1 * torch.randn(pos.shape, generator=g3, dtype=dtype)
        pos = pos + 0.02 * cfg3.dx * torch.randn(pos.shape, generator=g3, dtype=dtype)
        if name == "closure":
            # lift the top layers so their summation density falls below threshold
            top = pos[:, 1] > cfg3.column_height - 1.5 * cfg3.dx
            pos[top, 1] += 0.3 * cfg3.dx
        h = cfg3.h_factor * cfg3.dx
        allp = pos if wall is None else torch.cat((pos, wall))
        pairs = None
        if with_pairs:
            cache = PairListCache(2.0 * h, 0.25 * h, fluid_count=pos.shape[0])
            pairs, _ = cache.get(allp)
        rho, p, acc, reaction = db3.rhs_3d(pos, vel, mass, cfg3, pairs=pairs, wall=wall,
                                           return_wall_force=True)
        out.update({name + "_pos": _np(pos), name + "_vel": _np(vel), name + "_mass": _np(mass),
                    name + "_rho": _np(rho), name + "_p": _np(p), name + "_acc": _np(acc),
                    name + "_reaction": _np(reaction)})
        if wall is not None:
            out[name + "_wall"] = _np(wall)
        if pairs is not None:
            out[name + "_pairs"] = _np(pairs)
        out[name + "_params"] = np.array([cfg3.dx, cfg3.h_factor, cfg3.rho0, cfg3.sound_speed,
                                          cfg3.gamma, cfg3.gravity, cfg3.viscosity,
                                          cfg3.wall_friction, cfg3.dt, cfg3.density_diffusion,
                                          db3._mass_3d(cfg3.dx, cfg3.h_factor, cfg3.rho0, True)])
    torch.set_default_dtype(torch.float64)
    np.savez_compressed(HERE / "rhs3d_opts.npz", **out)

    # dam_break_3d.run_case on a small tank, fixed and adaptive step
    out = {}
    for name, over in (("fixed", {"dt": 1e-3}), ("adaptive", {"adaptive_dt": True, "dt": 1.25e-3})):
        cfg3 = db3.DamBreak3DConfig(dx=0.05, tank_height=0.5, tank_depth=0.25,
                                    column_width=0.2, column_depth=0.25, column_height=0.3,
                                    pool_width=0.6, viscosity=1e-3, sound_speed=15.0,
                                    steps=80, output_every=10, wall_friction=0.1, **over)
        run = db3.run_case(cfg3, snapshot_times=[0.03])
        out.update({name + "_times": np.array(run.times), name + "_front": np.array(run.front),
                    name + "_ke": np.array(run.kinetic_energy),
                    name + "_pe": np.array(run.potential_energy),
                    name + "_scalars": np.array([run.escaped, run.rebuilds, run.max_speed,
                                                 run.max_density_compression,
                                                 run.compression_before_wall_impact,
                                                 -1.0 if run.wall_impact_time is None
                                                 else run.wall_impact_time,
                                                 run.initial_energy, run.final_energy,
                                                 run.steps_taken, run.simulated_time, run.dt_min,
                                                 run.dt_max, run.max_acoustic_courant,
                                                 run.initial_front]),
                    name + "_final": _np(run.final_positions),
                    name + "_snap_t": np.array([s[0] for s in run.snapshots]),
                    name + "_snap_pos": np.stack([_np(s[1]) for s in run.snapshots])})
    np.savez_compressed(HERE / "dam3d_run.npz", **out)
    print("moving-wall and rhs_3d fixtures written to", HERE)


if __name__ == "__main__":
    args = sys.argv[1:]
    if args and args[0] == "--only-moving-walls":
        moving_walls(*args[1:])
    elif args and args[0] == "--only-shear":
        shear(*args[1:])
    elif args and args[0] == "--only-order4":
        order4(*args[1:])
    elif args and args[0] == "--only-rhs-options":
        rhs_options(*args[1:])
    elif args and args[0] == "--only-kleefsman":
        kleefsman_gauges(*args[1:])
    elif args and args[0] == "--only-shift":
        lind_shift(*args[1:])
    elif args and args[0] == "--only-tvc":
        tvc(*args[1:])
    elif args and args[0] == "--only-kleefsman-adaptive":
        kleefsman_adaptive(*args[1:])
    else:
        main(*args)
        shear(*args)
        order4(*args)
        rhs_options(*args)
        kleefsman_gauges(*args)
        lind_shift(*args)
        tvc(*args)
        kleefsman_adaptive(*args)
        moving_walls(*args)
