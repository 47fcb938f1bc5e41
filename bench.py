"""Benchmark: 2-D WCSPH dam break on one MI355X, FP32-compensated vs FP64.

Workload (the 400k-1.6M-particle dam break the metric is quoted on): a
1 x 2 m water column at the left of a 4 x 3 m tank with three wall layers on
a square lattice, h = 1.3 dx, hydrostatic initial masses, fluid at rest,
c0 = 10 sqrt(2 g H) with H = 2 m, dt at acoustic Courant 0.3, nu = 0.01,
no density diffusion, Verlet skin 0.25 h.  dx = sqrt(2 / N) with N = 1.6 M
gives 1,598,472 fluid + 26,847 wall particles.

One step = one full pass of the particle pipeline (density + Tait, Adami wall
pressure, pressure + viscous force, symplectic Euler + clamp, and the
amortised neighbour-list rebuilds) over all fluid particles.  The state is
resident in HBM when the timed region starts.

Prints one JSON line; ``value`` is fluid particle-updates/s of the
FP32-compensated pipeline.  ``roofline`` prices the force sweep (the
dominant kernel) at its algorithmic bytes over its HIP-event time;
``cpu_baseline`` times the FP64 CPU restatement on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
PROFILE = os.path.join(HERE, "profiles", "k_tile_force_dambreak_1p6M_fp32c.md")
PROFILE_LISTS = os.path.join(HERE, "profiles", "k_step_force_dambreak_1p6M_fp32c.md")
PROFILE_3D = os.path.join(HERE, "profiles", "k_step_force_kleefsman_5M_fp32c.md")


def profiled_traffic(nf: int, path: str = PROFILE, kernel: str = "k_tile_force"):
    """HBM bytes per force-sweep launch from a committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary, when it was taken on this workload; else None."""
    try:
        with open(path) as f:
            text = f.read()
    except OSError:
        return None
    if f"{nf:,} fluid" not in text:
        return None
    m = re.search(rf"traffic_bytes_per_launch {kernel} = (\d+)", text)
    return int(m.group(1)) if m else None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fluid", type=float, default=1.6e6, help="target fluid particle count")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--diffusion", type=float, default=0.0,
                    help="dam break with Molteni δ-diffusion of this δ (slot-list sweeps; "
                         "the LDS tile force sweep carries no diffusion term)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--tile", default=None,
                    help="fixed layout instead of the autotuner's choice: a tile shape TXxTY "
                         "(e.g. 8x6), 'lists' or 'halo'")
    ap.add_argument("--no-serve", action="store_true", help="skip the served δ⁺ batch")
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-steps", type=int, default=500)
    ap.add_argument("--serve-sequential", type=int, default=16,
                    help="requests of the stream timed alone on the batch engine")
    ap.add_argument("--serve-sequential-api", type=int, default=1,
                    help="1: also time the sample through the single-case API loop")
    ap.add_argument("--workload", choices=("dambreak", "shear", "kleefsman"), default="dambreak",
                    help="dambreak: the 2-D dam break line (default); shear: 3-D periodic shear "
                         "wave; kleefsman: 3-D dam break against a box obstacle")
    ap.add_argument("--shear-n", type=int, default=128, help="lattice edge of the shear-wave cube")
    ap.add_argument("--decay-n", type=int, default=64,
                    help="lattice edge of the full-length decay run of the shear workload (0: skip)")
    ap.add_argument("--fluid3d", type=float, default=5e6, help="fluid particles of the box case")
    ap.add_argument("--t-end", type=float, default=0.0,
                    help="box case: also integrate both precisions to this physical time and "
                         "compare the force / moment histories on the box")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates rocprofv3 FETCH/WRITE_SIZE)")
    return ap.parse_args()


def dam_break_state(n_target: float):
    from tensorsph_amd.config import DamBreakConfig
    from tensorsph_amd.cases import make_column, wall_lattice, lattice_mass_unit
    dx = math.sqrt(2.0 / n_target)
    c0 = 10.0 * math.sqrt(2.0 * 9.81 * 2.0)
    h = 1.3 * dx
    cfg = DamBreakConfig(dx=dx, tank_height=3.0, column_width=1.0, column_height=2.0,
                         pool_width=4.0, sound_speed=c0, dt=0.3 * h / c0, viscosity=0.01)
    pos, mass = make_column(cfg)
    wall = wall_lattice(cfg, top=cfg.tank_top())
    return cfg, pos, mass, wall, lattice_mass_unit(cfg)


def make_sim(cfg, pos, mass, wall, wmass, precision, diffusion=0.0):
    import numpy as np
    from tensorsph_amd.engine import WcsphSimulation
    h = cfg.h_factor * cfg.dx
    return WcsphSimulation(pos, np.zeros_like(pos), mass, wall, h=h, rho0=cfg.rho0,
                           sound_speed=cfg.sound_speed, gamma=cfg.gamma, gravity=cfg.gravity,
                           viscosity=cfg.viscosity, dt=cfg.dt, wall_mass=wmass,
                           skin=cfg.pair_skin_factor * h, lo=[0.0, 0.0],
                           hi=[cfg.pool_width, cfg.tank_top()], tol=0.25 * cfg.dx,
                           precision=precision, density_diffusion=diffusion)


def time_gpu(sim, steps, warmup, tile=None):
    import torch
    from tensorsph_amd.autotune import tune_tiles
    if tile in ("lists", "halo"):
        sim.configure(tiles=False, halo=tile == "halo")
        tuned = {"best": tile, "ms": {}}
        sim.step(max(warmup, 1))
    elif tile is not None:
        sim.configure(tiles=True, tile=tile)
        tuned = {"best": tuple(tile), "ms": {}}
        sim.step(max(warmup, 1))
    else:
        sim.step(max(warmup // 2, 1))
        tuned = tune_tiles(sim, steps=4, warm=1)     # trial steps count as warm-up
        sim.step(max(warmup - warmup // 2, 1))
    torch.cuda.synchronize()
    sim.set_timing(True)
    t0 = time.perf_counter()
    sim.step(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sweep_ms, nsteps = sim.sweep_times()
    sim.set_timing(False)
    return wall, sweep_ms, nsteps, tuned


def cpu_baseline(seconds: float):
    """FP64 oracle loop (C, one thread) on the same dam break at a 25k-particle sample."""
    import numpy as np
    from oracle import oracle as orc
    cfg, pos, mass, wall, wmass = dam_break_state(2.5e4)
    h = cfg.h_factor * cfg.dx
    st = orc.Stepper2D(pos, np.zeros_like(pos), mass, wall, h=h, rho0=cfg.rho0,
                       c0=cfg.sound_speed, gamma=cfg.gamma, gravity=cfg.gravity,
                       nu=cfg.viscosity, dt=cfg.dt, wall_mass=wmass,
                       skin=cfg.pair_skin_factor * h, lo=[0.0, 0.0],
                       hi=[cfg.pool_width, cfg.tank_top()], tol=0.25 * cfg.dx)
    st.step()                      # first list build outside the timing
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < seconds:
        st.step()
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": pos.shape[0] * steps / dt, "unit": "particle-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"FP64 C restatement, 1 thread, dam break at {pos.shape[0]} fluid + "
                      f"{wall.shape[0]} wall particles (dx={cfg.dx:.4g}), {steps} steps in "
                      f"{dt:.1f} s, list rebuilds included"}


def serving_bench(n_requests: int, steps: int, n_seq: int, n_api: int) -> dict:
    """The served δ⁺ workload: Poisson requests of 2k-250k-particle drop / impact
    runs, continuous batching on one engine holding ~40 % of the stream's rows,
    against sequential single-case runs of samples of the same requests: alone
    on the batch engine, and through the single-case API loop of the
    reference's drivers (dt, pair cache, delta_plus_step per step)."""
    from tensorsph_amd.serving import (BatchEngine, build_case, request_stream, run_sequential,
                                       run_sequential_api, serve)
    reqs = request_stream(count=n_requests, seed=512, steps=steps)
    cases = [build_case(k, kw) for k, kw in reqs]
    rows = sum(c.n for c in cases)
    walls = sum(c.n_wall for c in cases)
    fluid_cap = int(0.4 * rows) + max(c.n for c in cases)
    wall_cap = int(0.4 * walls) + max(c.n_wall for c in cases)
    eng = BatchEngine(fluid_cap, wall_cap, sims=256)
    eng.step(0)                                   # capture the step graphs
    footprint = eng.footprint_bytes()
    rate = n_requests / 20.0                      # requests arrive faster than they finish
    rep = serve(eng, cases, rate)
    eng.close()
    del cases
    summary = rep.summary()
    stride = max(1, n_requests // max(n_seq, 1))
    sample = reqs[::stride][:n_seq]
    seq_s, _ = run_sequential(sample)
    seq_rate = len(sample) / seq_s
    api = None
    if n_api:
        api_s, _ = run_sequential_api(sample)
        api_rate = len(sample) / api_s
        api = {"sims_per_s": api_rate, "seconds": api_s}
    return {
        "workload": f"{n_requests} δ⁺ drop / impact requests, {steps} steps each, "
                    f"{rows} fluid + {walls} wall particles in total, Poisson arrivals at "
                    f"{rate:.1f}/s, FP64",
        "sims_per_s": summary["sims_per_s"],
        "step_ms_p50": summary["step_ms_p50"], "step_ms_p99": summary["step_ms_p99"],
        "completion_s_p50": summary["completion_s_p50"], "completion_s_p99": summary["completion_s_p99"],
        "particle_updates_per_s": summary["particle_updates_per_s"],
        "batch_steps": summary["batch_steps"], "failed": summary["failed"],
        "engine_capacity": {"fluid_rows": fluid_cap, "wall_rows": wall_cap, "bytes": footprint,
                            "fraction_of_stream_rows": fluid_cap / rows},
        "sequential": {
            "sample": f"{len(sample)} requests of the stream (every {stride}th), one after another",
            "engine_alone": {"sims_per_s": seq_rate, "seconds": seq_s,
                             "how": "each case alone on a batch engine sized for it, step graphs "
                                    "captured before timing"},
            "single_case_api": api and dict(api, how="the reference drivers' loop per case: "
                                            "adaptive dt from max|v|, PairListCache(4 dx, 0.4 dx), "
                                            "delta_plus_step every step"),
        },
        "batched_over_engine_alone": summary["sims_per_s"] / seq_rate,
        "batched_over_single_case_api": summary["sims_per_s"] / api["sims_per_s"] if api else None,
    }


def _timed(step, steps, warmup):
    import torch
    step(max(warmup, 1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step(steps)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def _line(metric, value, steps, warmup, wall_s, dtype, config, **extra):
    out = {"metric": metric, "value": value, "unit": "particle-updates/s", "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": 1000.0 * wall_s / steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic: deterministic lattice (no dataset)", "config": config}
    out.update(extra)
    return out


def shear_bench(args) -> dict:
    """3-D triply periodic shear wave on the resident leapfrog stepper: unit
    cube, n^3 lattice jittered by U(-0.01 dx, 0.01 dx) from PCG64(20261203),
    h = 1.3 dx, rho0 = 1, nu = 0.01, c0 = 10, u_x = 1e-3 sin(2 pi z),
    renormalised Laplacian, dt = min(0.25 h / c0, 0.125 h^2 / nu), skin 0.5 h."""
    import numpy as np
    import torch
    from tensorsph_amd.threed.kernel3 import calibrated_particle_mass
    from tensorsph_amd.threed.solver import PeriodicSimulation, cubic_lattice, shear_wave_velocity
    n = args.shear_n
    dx = 1.0 / n
    h = 1.3 * dx
    nu, c0, rho0 = 0.01, 10.0, 1.0
    dt = min(0.25 * h / c0, 0.125 * h * h / nu)
    pos = cubic_lattice((n, n, n), dx)
    rng = np.random.Generator(np.random.PCG64(20261203))
    pos = (pos + torch.from_numpy(rng.uniform(-0.01 * dx, 0.01 * dx, size=tuple(pos.shape)))).remainder(1.0)
    vel = shear_wave_velocity(pos, 1e-3, 2.0 * math.pi, "single")
    mass = torch.full((pos.shape[0],), calibrated_particle_mass(dx, h, rho0), dtype=torch.float64)
    runs = {}
    for prec in ("fp32", "fp64"):
        sim = PeriodicSimulation(pos, vel, mass, h=h, box=1.0, rho0=rho0, sound_speed=c0, viscosity=nu,
                                 dt=dt, skin=0.5 * h, renormalise=True, precision=prec, device="cuda")
        wall_s = _timed(sim.step, args.steps, args.warmup)
        st = sim.stats()
        sim.close()
        runs[prec] = {"wall_s": wall_s, "value": pos.shape[0] * args.steps / wall_s, "pairs": st["pairs"],
                      "rebuilds": st["rebuilds"]}
    npart = pos.shape[0]
    v32, v64 = runs["fp32"]["value"], runs["fp64"]["value"]
    # compulsory bytes per particle-update: density sweep (x, m; rho written), the
    # moment pre-pass (x, m, rho; K written) and the force sweep + leapfrog
    # (x, v, m, rho, p, K; x, v written): 40 D + 4 s, plus 8 D + 2 s for K
    b32, b64 = 40 * 3 + 4 * 4 + 24 + 2 * 4, 40 * 3 + 4 * 8 + 24 + 2 * 8
    cfg = {"workload": f"3-D periodic shear wave, {n}^3 = {npart} particles, dx=1/{n}, h=1.3dx, "
                       f"rho0=1, nu=0.01, c0=10, dt={dt:.4g} s, jitter 0.01 dx",
           "precision": "pair arithmetic FP32 with Kahan sums (density FP64); positions, velocities "
                        "and leapfrog FP64", "particles": npart, "parallelism": "single GPU"}
    decay = shear_decay(args.decay_n) if args.decay_n > 0 else None
    return _line("particle-updates/s (3-D periodic shear wave, FP32-compensated) + HBM GB/s", v32,
                 args.steps, args.warmup, runs["fp32"]["wall_s"], "f32", cfg,
                 algorithmic_GBps_step=v32 * b32 / 1e9, algorithmic_bytes_per_update=b32,
                 pair_interactions_per_s=runs["fp32"]["pairs"] * args.steps / runs["fp32"]["wall_s"],
                 list_pairs=runs["fp32"]["pairs"], list_rebuilds_in_run=runs["fp32"]["rebuilds"],
                 fp64={"value": v64, "ms_per_step": 1000.0 * runs["fp64"]["wall_s"] / args.steps,
                       "algorithmic_GBps_step": v64 * b64 / 1e9},
                 fp32c_over_fp64=v32 / v64, decay=decay)


def shear_decay(n: int) -> dict:
    """The decay check of the shear workload: u_x = 1e-3 sin(2 pi z) on an n^3
    lattice (jitter 0.01 dx, PCG64(20261203)), nu = 0.01, rho0 = 1, c0 = 10,
    integrated to t = 1 / (4 pi^2 nu) with amplitude samples every 50 steps,
    against the analytic rate nu k^2 = 4 pi^2 nu of exp(-nu k^2 t)."""
    from tensorsph_amd.threed.solver import run_shear_wave
    nu = 0.01
    t_end = 1.0 / (4.0 * math.pi ** 2 * nu)
    out = {"lattice": f"{n}^3", "t_end_s": t_end}
    for name, dtype in (("fp32c", "float32"), ("fp64", "float64")):
        import torch
        t0 = time.perf_counter()
        res = run_shear_wave(n=n, nu=nu, amplitude=1e-3, sound_speed=10.0, rho0=1.0, t_end=t_end,
                             sample_every=50, mode="single", jitter=0.01, device="cuda",
                             dtype=getattr(torch, dtype))
        out[name] = {"fit_rate": res.fit_rate, "analytic_rate": res.analytic_rate,
                     "rel_error": res.rate_error, "steps": res.steps, "seconds": time.perf_counter() - t0,
                     "amplitude_final_over_initial": res.amplitudes[-1] / res.amplitudes[0]}
    return out


def kleefsman_bench(args) -> dict:
    """3-D dam break against a box obstacle (Kleefsman): 3.22 x 1 x 1 m tank,
    1.228 x 0.55 x 1 m column, 0.161 x 0.161 x 0.403 m box at x = 2.3955 m,
    3 wall layers on the tank, the box as a 3-layer shell, c0 = 20 m/s,
    dt = 0.9 of the acoustic Courant limit, h = 1.3 dx; the box force and moment
    are reduced every step."""
    import torch
    from tensorsph_amd.config import Kleefsman3DConfig
    from tensorsph_amd.kleefsman import dx_for_fluid, make_simulation
    cfg = Kleefsman3DConfig(dx=dx_for_fluid(args.fluid3d), sound_speed=20.0, wall_layers=3,
                            obstacle_fill="shell")
    cfg.dt = 0.9 * cfg.dt_from_courant()
    runs = {}
    for prec in ("fp32", "fp64"):
        sim, nf, nw = make_simulation(cfg, prec)
        sim.step(max(args.warmup, 1))
        torch.cuda.synchronize()
        sim.set_timing(True)
        t0 = time.perf_counter()
        sim.step(args.steps)
        torch.cuda.synchronize()
        wall_s = time.perf_counter() - t0
        sweep_ms, nsteps = sim.sweep_times()
        sim.set_timing(False)
        loads = sim.loads()
        stats = sim.stats()
        sim.close()
        runs[prec] = {"wall_s": wall_s, "value": nf * args.steps / wall_s,
                      "sweep_ms": [x / max(nsteps, 1) for x in sweep_ms], "rebuilds": stats["rebuilds"],
                      "force": loads[-1, 1:4].tolist(), "moment": loads[-1, 4:7].tolist()}
    v32, v64 = runs["fp32"]["value"], runs["fp64"]["value"]
    history = kleefsman_history(cfg, args.t_end) if args.t_end > 0 else None
    s = 4
    b_force = nf * (48 * 2 + 2 * s) + nw * (24 + 3 * s)
    force_ms = runs["fp32"]["sweep_ms"][2]
    achieved = b_force / (force_ms * 1e-3) / 1e9 if force_ms > 0 else 0.0
    f32, f64 = [runs[p]["force"] for p in ("fp32", "fp64")]
    scale = max(abs(x) for x in f64) or 1.0
    cfg_out = {"workload": f"3-D dam break with a box obstacle, {nf} fluid + {nw} wall particles, "
                           f"dx={cfg.dx:.5g} m, h=1.3dx, c0=20 m/s, dt={cfg.dt:.4g} s",
               "precision": "pair arithmetic FP32 with Kahan sums (density FP64); positions, "
                            "velocities and integration FP64", "fluid": nf, "wall": nw,
               "parallelism": "single GPU"}
    return _line("particle-updates/s (3-D dam break with a box obstacle, FP32-compensated) + HBM GB/s",
                 v32, args.steps, args.warmup, runs["fp32"]["wall_s"], "f32", cfg_out,
                 sweep_ms_per_step=dict(zip(("density", "wall", "force_integrate"), runs["fp32"]["sweep_ms"])),
                 roofline={"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS,
                           "traffic": profiled_traffic(nf, PROFILE_3D, "k_step_force"),
                           "traffic_source": os.path.relpath(PROFILE_3D, HERE),
                           "kernel": "k_step_force (force + viscosity + integrate + box clamp)",
                           "algorithmic_bytes_per_launch": b_force},
                 list_rebuilds_in_run=runs["fp32"]["rebuilds"],
                 box_load_last_step={"force_N_fp32c": f32, "force_N_fp64": f64,
                                     "moment_Nm_fp32c": runs["fp32"]["moment"],
                                     "moment_Nm_fp64": runs["fp64"]["moment"],
                                     "force_rel_diff": max(abs(a - b) for a, b in zip(f32, f64)) / scale},
                 fp64={"value": v64, "ms_per_step": 1000.0 * runs["fp64"]["wall_s"] / args.steps,
                       "sweep_ms_per_step": dict(zip(("density", "wall", "force_integrate"),
                                                     runs["fp64"]["sweep_ms"]))},
                 fp32c_over_fp64=v32 / v64, history=history)


def kleefsman_history(cfg, t_end: float) -> dict:
    """Both precisions integrated to t_end with the box force and moment of
    every step (kleefsman_obstacle_3d.py:644-706): peak force, impulse, first
    contact and the FP32c - FP64 difference of the histories."""
    import numpy as np
    import torch
    from tensorsph_amd.kleefsman import make_simulation
    steps = int(math.ceil(t_end / cfg.dt))
    loads = {}
    nf = 0
    for prec in ("fp32", "fp64"):
        sim, nf, _ = make_simulation(cfg, prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < steps:                     # chunks, with a progress line after each
            k = min(1000, steps - done)
            sim.step(k)
            done += k
            print(f"box history {prec}: step {done}/{steps}, {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        loads[prec] = (sim.loads(), secs, sim.stats())
        sim.close()
    out = {"t_end_s": steps * cfg.dt, "steps": steps}
    f64 = loads["fp64"][0][:, 1:4]
    scale = float(np.abs(f64).max()) or 1.0
    for prec, (rows, secs, st) in loads.items():
        f, m = rows[:, 1:4], rows[:, 4:7]
        mag = np.sqrt((f * f).sum(1))
        hit = np.nonzero(mag > 1e-3 * max(mag.max(), 1e-30))[0]
        out[prec] = {"seconds": secs, "particle_updates_per_s": nf * steps / secs,
                     "peak_force_N": f[int(np.argmax(mag))].tolist(), "peak_force_time_s": float(np.argmax(mag) * cfg.dt),
                     "impulse_Ns": (f.sum(0) * cfg.dt).tolist(), "peak_moment_Nm": float(np.sqrt((m * m).sum(1)).max()),
                     "first_contact_s": float(hit[0] * cfg.dt) if hit.size else None,
                     "clamped_particles": int(rows[-1, 10]), "rebuilds": int(st["rebuilds"])}
    out["force_history_max_diff_over_peak"] = float(np.abs(loads["fp32"][0][:, 1:4] - f64).max() / scale)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 (a slab "
                                   "decomposition over several GPUs is designed, not built)"}))
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    if args.copy_probe:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 8}))
        return
    if args.workload == "shear":
        print(json.dumps(shear_bench(args)))
        return
    if args.workload == "kleefsman":
        print(json.dumps(kleefsman_bench(args)))
        return
    cfg, pos, mass, wall, wmass = dam_break_state(args.fluid)
    nf, nw = pos.shape[0], wall.shape[0]
    h = cfg.h_factor * cfg.dx

    if args.tile is None or args.tile.lower() in ("lists", "halo"):
        tile = args.tile and args.tile.lower()
    else:
        tile = tuple(int(v) for v in args.tile.lower().split("x"))
    if args.diffusion > 0.0 and tile != "halo":
        tile = "lists"
    sim = make_sim(cfg, pos, mass, wall, wmass, "fp32", args.diffusion)
    wall_s, sweep_ms, nsteps, tuned = time_gpu(sim, args.steps, args.warmup, tile)
    pairs = sim.count_pairs(2.0 * h)
    stats = sim.stats()
    sim.close()
    del sim
    value = nf * args.steps / wall_s
    ms_step = 1000.0 * wall_s / args.steps

    # algorithmic bytes per launch of the force sweep (sweep B): reads x, v (FP64,
    # 2 x 16 B), m, rho (4 B each), writes x, v (2 x 16 B) per fluid particle,
    # plus each wall particle's x, m, rho, p read once (16 + 3 x 4 B)
    s = 4
    b_force = nf * (32 * 2 + 2 * s) + nw * (16 + 3 * s)
    b_density = nf * (16 + 2 * s) + nw * (16 + s)
    force_ms = sweep_ms[2] / max(nsteps, 1)
    on_lists = args.diffusion > 0.0 or tuned["best"] in ("lists", "halo")
    kernel, prof = ("k_step_force", PROFILE_LISTS) if on_lists else ("k_tile_force", PROFILE)
    achieved = b_force / (force_ms * 1e-3) / 1e9 if force_ms > 0 else 0.0
    step_bytes = nf * (40 * 2 + 4 * s) + nw * (24 * 2 + 2 * s)

    result = {
        "metric": "particle-updates/s (2-D WCSPH dam break, FP32-compensated) + HBM GB/s",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: deterministic dam-break lattice (no dataset)",
        "config": {"workload": f"2-D dam break, {nf} fluid + {nw} wall particles, "
                               f"dx={cfg.dx:.5g} m, h=1.3dx, c0={cfg.sound_speed:.4g} m/s, "
                               f"dt={cfg.dt:.4g} s, nu={cfg.viscosity}"
                               + (f", Molteni δ-diffusion δ={args.diffusion}" if args.diffusion else ""),
                   "precision": "pair arithmetic FP32 with Kahan sums; positions, velocities "
                                "and integration FP64",
                   "fluid": nf, "wall": nw, "parallelism": "single GPU"},
        "pair_interactions_per_s": pairs * args.steps / wall_s,
        "pairs_within_2h": pairs,
        "list_rebuilds_in_run": stats["rebuilds"],
        "tile_shape": {"chosen": tuned["best"], "trial_ms_per_step": tuned["ms"]},
        "algorithmic_GBps_step": step_bytes * args.steps / wall_s / 1e9,
        "sweep_ms_per_step": {"density": sweep_ms[0] / max(nsteps, 1),
                              "wall": sweep_ms[1] / max(nsteps, 1),
                              "force_integrate": force_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if args.diffusion else profiled_traffic(nf, prof, kernel),
                     "traffic_source": None if args.diffusion else os.path.relpath(prof, HERE),
                     "kernel": (f"{kernel} (slot lists: force + viscosity + integrate)" if on_lists
                                else f"{kernel} (LDS tiles: force + viscosity + integrate)"),
                     "algorithmic_bytes_per_launch": b_force},
    }
    if not args.no_fp64:
        sim = make_sim(cfg, pos, mass, wall, wmass, "fp64", args.diffusion)
        w64, sw64, n64, tuned64 = time_gpu(sim, args.steps, args.warmup, tile)
        sim.close()
        del sim
        v64 = nf * args.steps / w64
        f64_ms = sw64[2] / max(n64, 1)
        b64 = nf * (32 * 2 + 2 * 8) + nw * (16 + 3 * 8)
        result["fp64"] = {"value": v64, "ms_per_step": 1000.0 * w64 / args.steps,
                          "force_sweep_ms": f64_ms,
                          "force_sweep_GBps": b64 / (f64_ms * 1e-3) / 1e9 if f64_ms > 0 else 0.0}
        result["fp32c_over_fp64"] = value / v64
    if not args.no_serve:
        result["serving"] = serving_bench(args.serve_requests, args.serve_steps, args.serve_sequential,
                                          args.serve_sequential_api)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
