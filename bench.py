"""Benchmark: 2-D WCSPH dam break on one MI355X, FP32-compensated vs FP64.

Workload (the 400k-1.6M-particle dam break the metric is quoted on): a
1 x 2 m water column at the left of a 4 x 3 m tank with three wall layers on
a square lattice, h = 1.3 dx, hydrostatic initial masses, fluid at rest,
c0 = 10 sqrt(2 g H) with H = 2 m, dt at acoustic Courant 0.3, nu = 0.01,
no density diffusion, Verlet skin 0.25 h.  dx = sqrt(2 / N) with N = 1.6 M
gives 1,598,472 fluid + 26,847 wall particles.

The flow is first developed untimed (the collapse runs to t = --develop-time,
about 0.25 s, where the front moves at several m/s and the lists go stale
every few steps); the timed window of --steps steps then includes every list
rebuild that falls in it.  One step = one full pass of the particle pipeline
(density + Tait, Adami wall pressure, pressure + viscous force, symplectic
Euler + clamp) over all fluid particles, plus the rebuilds.  The state is
resident in HBM when the timed region starts.

Prints one JSON line; ``value`` is fluid particle-updates/s of the
FP32-compensated pipeline in the window.  ``roofline`` prices the force
sweep (the dominant kernel) at its algorithmic bytes over its mean HIP-event
duration, sampled on every 10th step of a second window; ``cpu_baseline``
times the multi-threaded CPU port (oracle/tsph_cpu_omp.c) from the same
developed state on the host's cores and on one thread.  Blocks for the 3-D
shear wave (192^3) and the 3-D box case (20 M), the FP32c-vs-FP64 error on a
common state and the served delta-plus batch ride on the same line.
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
# δ⁺ step (FP64), compulsory bytes per particle-update: the state x, v (2 x 16 B),
# rho, m (2 x 8 B) read and x, v, rho (40 B) written, plus the step's
# intermediates written once and read once (new velocity 16 B, lambda 8 B,
# the renormalisation matrix 32 B, normal 16 B, band flags 8 B): 248 B, taken
# as 250 B
DP_BYTES_PER_UPDATE = 250
PROFILE = os.path.join(HERE, "profiles", "k_tile_force_dambreak_1p6M_fp32c.md")
PROFILE_LISTS = os.path.join(HERE, "profiles", "k_step_force_dambreak_1p6M_fp32c.md")
PROFILE_3D = os.path.join(HERE, "profiles", "k_step_force_halo_kleefsman_20M_fp32c.md")
PROFILE_HALO = os.path.join(HERE, "profiles", "k_step_force_halo_dambreak_1p6M_fp32c.md")


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
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed steps after the flow is developed (and the layout chosen)")
    ap.add_argument("--develop-time", type=float, default=0.25,
                    help="dam break: physical time run untimed before the window")
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
    ap.add_argument("--no-3d", action="store_true", help="skip the 3-D blocks of the dam-break line")
    ap.add_argument("--block-steps", type=int, default=40, help="timed steps of each 3-D block")
    ap.add_argument("--block-warmup", type=int, default=10)
    ap.add_argument("--shear-n", type=int, default=192, help="lattice edge of the shear-wave cube")
    ap.add_argument("--decay-n", type=int, default=64,
                    help="lattice edge of the full-length decay run of the shear workload (0: skip)")
    ap.add_argument("--fluid3d", type=float, default=20e6, help="fluid particles of the box case")
    ap.add_argument("--layout3d", choices=("default", "lists", "halo"), default="default",
                    help="sweep layout of the box case (default: the stepper's own)")
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


def make_sim(cfg, pos, mass, wall, wmass, precision, diffusion=0.0, vel=None):
    import numpy as np
    from tensorsph_amd.engine import WcsphSimulation
    h = cfg.h_factor * cfg.dx
    return WcsphSimulation(pos, np.zeros_like(pos) if vel is None else vel, mass, wall, h=h, rho0=cfg.rho0,
                           sound_speed=cfg.sound_speed, gamma=cfg.gamma, gravity=cfg.gravity,
                           viscosity=cfg.viscosity, dt=cfg.dt, wall_mass=wmass,
                           skin=cfg.pair_skin_factor * h, lo=[0.0, 0.0],
                           hi=[cfg.pool_width, cfg.tank_top()], tol=0.25 * cfg.dx,
                           precision=precision, density_diffusion=diffusion)


def use_layout(sim, layout):
    if layout in ("lists", "halo"):
        sim.configure(tiles=False, halo=layout == "halo")
    else:
        sim.configure(tiles=True, tile=layout)


def develop(sim, steps: int, label: str) -> None:
    """Untimed steps, in chunks with a progress line after each."""
    import torch
    done, t0 = 0, time.perf_counter()
    while done < steps:
        k = min(5000, steps - done)
        sim.step(k)
        done += k
        torch.cuda.synchronize()
        print(f"{label}: developed {done}/{steps} steps, {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)


def timed_window(sim, steps: int):
    """Wall time of ``steps`` steps (list rebuilds included) and the rebuilds in them."""
    import torch
    torch.cuda.synchronize()
    b0 = sim.stats()["rebuilds"]
    t0 = time.perf_counter()
    sim.step(steps)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, sim.stats()["rebuilds"] - b0


def kernel_window(sim, steps: int, every: int = 10):
    """Mean HIP-event durations (ms) of the density, wall and force sweeps over
    every ``every``-th step of a window of ``steps`` steps, and the count."""
    import torch
    sim.set_timing(True, every)
    sim.step(steps)
    torch.cuda.synchronize()
    ms, n = sim.sweep_times()
    sim.set_timing(False)
    return [x / max(n, 1) for x in ms], n


def precision_error(cfg, mass, wall, wmass, state, diffusion=0.0) -> dict:
    """FP32-compensated against FP64 on one common state (the developed FP64
    state): one step from it in each precision; the density and acceleration
    of that step's evaluation, largest difference over the FP64 field's
    largest magnitude."""
    import numpy as np
    pos, vel = state
    out = {}
    for prec in ("fp32", "fp64"):
        sim = make_sim(cfg, pos, mass, wall, wmass, prec, diffusion, vel=vel)
        try:
            sim.step(1)
            r = sim.read()
            out[prec] = {k: r[k].double().cpu().numpy() for k in ("density", "acceleration")}
        finally:
            sim.close()

    def scaled(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))

    return {"density_max_rel": scaled(out["fp32"]["density"], out["fp64"]["density"]),
            "acceleration_max_over_max": scaled(out["fp32"]["acceleration"], out["fp64"]["acceleration"]),
            "state": "the FP64 run's state after its timed window; one step in each precision"}


def common_state_error(make, state) -> dict:
    """FP32-compensated against FP64 from one common state: make(prec, pos,
    vel) builds a simulation there; one step in each precision, then the
    largest difference of density and acceleration over the FP64 field's
    largest magnitude."""
    import numpy as np
    out = {}
    for prec in ("fp32", "fp64"):
        sim = make(prec, *state)
        try:
            sim.step(1)
            r = sim.read()
            out[prec] = {k: r[k].double().cpu().numpy() for k in ("density", "acceleration")}
        finally:
            sim.close()

    def scaled(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))

    return {"density_max_rel": scaled(out["fp32"]["density"], out["fp64"]["density"]),
            "acceleration_max_over_max": scaled(out["fp32"]["acceleration"], out["fp64"]["acceleration"])}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(cap)) if cap and cap.isdigit() else n


def cpu_baseline(cfg, wmass, wall, state, seconds: float) -> dict:
    """The multi-threaded CPU port of the same FP64 step (oracle/tsph_cpu_omp.c,
    pinned against the oracle's Stepper2D), started from the GPU run's
    developed state: all host threads for ``seconds``, then one thread for
    half as long.  The first list build is outside the timing; later ones
    are inside."""
    import numpy as np
    from oracle import oracle as orc
    h = cfg.h_factor * cfg.dx
    pos, vel, mass = state
    kw = dict(h=h, rho0=cfg.rho0, c0=cfg.sound_speed, gamma=cfg.gamma, gravity=cfg.gravity,
              nu=cfg.viscosity, dt=cfg.dt, wall_mass=wmass, skin=cfg.pair_skin_factor * h,
              lo=[0.0, 0.0], hi=[cfg.pool_width, cfg.tank_top()], tol=0.25 * cfg.dx)
    out = {}
    for label, threads, budget in (("all", host_threads(), seconds), ("one", 1, 0.5 * seconds)):
        st = orc.OmpStepper2D(pos, vel, mass, wall, threads=threads, **kw)
        try:
            st.step(1)                 # first list build
            t0 = time.perf_counter()
            steps = 0
            while time.perf_counter() - t0 < budget:
                st.step(1)
                steps += 1
            dt = time.perf_counter() - t0
            rebuilds = st.read()["rebuilds"] - 1
        finally:
            st.close()
        out[label] = {"value": pos.shape[0] * steps / dt, "threads": st.threads, "steps": steps,
                      "seconds": dt, "list_rebuilds": rebuilds}
    return {"value": out["all"]["value"], "unit": "particle-updates/s", "cores": out["all"]["threads"],
            "kind": "port", "cpu_model": cpu_model(),
            "one_thread": {"value": out["one"]["value"], "steps": out["one"]["steps"],
                           "seconds": out["one"]["seconds"]},
            "sample": f"FP64 OpenMP port of the step (oracle/tsph_cpu_omp.c), the same dam break at "
                      f"{pos.shape[0]} fluid + {wall.shape[0]} wall particles from the GPU run's "
                      f"developed state: {out['all']['steps']} steps in {out['all']['seconds']:.1f} s on "
                      f"{out['all']['threads']} threads ({out['all']['list_rebuilds']} list rebuilds), "
                      f"{out['one']['steps']} steps in {out['one']['seconds']:.1f} s on 1 thread"}


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
    rep = serve(eng, cases, rate, tune=True)
    launch = eng.launch
    eng.close()
    del cases
    summary = rep.summary()
    stride = max(1, n_requests // max(n_seq, 1))
    sample = reqs[::stride][:n_seq]
    seq_s, _ = run_sequential(sample, launch=launch)      # the tuned launch shape too
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
        "autotune": summary["autotune"],
        "engine_capacity": {"fluid_rows": fluid_cap, "wall_rows": wall_cap, "bytes": footprint,
                            "fraction_of_stream_rows": fluid_cap / rows},
        "sequential": {
            "sample": f"{len(sample)} requests of the stream (every {stride}th), one after another",
            "engine_alone": {"sims_per_s": seq_rate, "seconds": seq_s,
                             "how": "each case alone on a batch engine sized for it, with the "
                                    "launch shape the served engine's autotuner chose, step graphs "
                                    "captured before timing"},
            "single_case_api": api and dict(api, how="the reference drivers' loop per case: "
                                            "adaptive dt from max|v|, PairListCache(4 dx, 0.4 dx), "
                                            "delta_plus_step every step"),
        },
        "roofline": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                     "algorithmic_bytes_per_update": DP_BYTES_PER_UPDATE,
                     "achieved": summary["particle_updates_per_s"] * DP_BYTES_PER_UPDATE / 1e9,
                     "frac": summary["particle_updates_per_s"] * DP_BYTES_PER_UPDATE / 1e9 / HBM_PEAK_GBS,
                     "how": "served particle-updates/s times the frozen FP64 δ⁺ bytes per update"},
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
    make = lambda prec, p0, v0: PeriodicSimulation(p0, v0, mass, h=h, box=1.0, rho0=rho0, sound_speed=c0,
                                                   viscosity=nu, dt=dt, skin=0.5 * h, renormalise=True,
                                                   precision=prec, device="cuda")
    state = None
    for prec in ("fp32", "fp64"):
        sim = make(prec, pos, vel)
        wall_s = _timed(sim.step, args.steps, args.warmup)
        st = sim.stats()
        if prec == "fp64":
            r = sim.read()
            state = (r["position"].cpu(), r["velocity"].cpu())
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
                 fp32c_over_fp64=v32 / v64,
                 fp32c_vs_fp64_error=dict(common_state_error(make, state),
                                          state="the FP64 run's state after its timed window; one "
                                                "step in each precision"),
                 decay=decay)


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
    state = None
    layout = getattr(args, "layout3d", "default")
    for prec in ("fp32", "fp64"):
        sim, nf, nw = make_simulation(cfg, prec)
        if layout != "default":
            sim.configure(tiles=False, halo=layout == "halo")
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
        lay = sim.layout()
        if prec == "fp64":
            r = sim.read()
            state = (r["position"].cpu().numpy(), r["velocity"].cpu().numpy())
        sim.close()
        runs[prec] = {"wall_s": wall_s, "value": nf * args.steps / wall_s,
                      "sweep_ms": [x / max(nsteps, 1) for x in sweep_ms], "rebuilds": stats["rebuilds"],
                      "force": loads[-1, 1:4].tolist(), "moment": loads[-1, 4:7].tolist(), "layout": lay}
    v32, v64 = runs["fp32"]["value"], runs["fp64"]["value"]
    history = kleefsman_history(cfg, args.t_end, layout) if args.t_end > 0 else None
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
                 layout={"fp32c": runs["fp32"]["layout"], "fp64": runs["fp64"]["layout"]},
                 box_load_last_step={"force_N_fp32c": f32, "force_N_fp64": f64,
                                     "moment_Nm_fp32c": runs["fp32"]["moment"],
                                     "moment_Nm_fp64": runs["fp64"]["moment"],
                                     "force_rel_diff": max(abs(a - b) for a, b in zip(f32, f64)) / scale},
                 fp64={"value": v64, "ms_per_step": 1000.0 * runs["fp64"]["wall_s"] / args.steps,
                       "sweep_ms_per_step": dict(zip(("density", "wall", "force_integrate"),
                                                     runs["fp64"]["sweep_ms"]))},
                 fp32c_over_fp64=v32 / v64,
                 fp32c_vs_fp64_error=dict(common_state_error(
                     lambda prec, p0, v0: make_simulation(cfg, prec, pos=p0, vel=v0)[0], state),
                     state="the FP64 run's state after its timed window; one step in each precision"),
                 history=history)


def kleefsman_history(cfg, t_end: float, layout: str = "default") -> dict:
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
        if layout != "default":
            sim.configure(tiles=False, halo=layout == "halo")
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
    print(json.dumps(dam_break_bench(args)))


def dam_break_bench(args) -> dict:
    import numpy as np
    from tensorsph_amd.autotune import tune_tiles
    cfg, pos, mass, wall, wmass = dam_break_state(args.fluid)
    nf, nw = pos.shape[0], wall.shape[0]
    h = cfg.h_factor * cfg.dx
    dev_steps = int(round(args.develop_time / cfg.dt))
    if args.tile is None or args.tile.lower() in ("lists", "halo"):
        layout = args.tile and args.tile.lower()
    else:
        layout = tuple(int(v) for v in args.tile.lower().split("x"))
    if args.diffusion > 0.0 and layout not in (None, "lists", "halo"):
        layout = "halo"
    runs = {}
    tuned = {"best": layout, "ms": {}}
    for prec in ("fp32",) + (() if args.no_fp64 else ("fp64",)):
        sim = make_sim(cfg, pos, mass, wall, wmass, prec, args.diffusion)
        try:
            develop(sim, dev_steps, f"dam break {prec}")
            warm = dev_steps
            if prec == "fp32" and layout is None:
                # the layout tried on the developed flow (trial steps advance it)
                tuned = tune_tiles(sim, candidates=("halo", "lists"), steps=20, warm=2)
                warm += 2 * 22
            else:
                use_layout(sim, tuned["best"])
            sim.step(max(args.warmup, 1))
            warm += max(args.warmup, 1)
            t_start = float(sim.time_state()["elapsed"])
            wall_s, rebuilds = timed_window(sim, args.steps)
            sweep, ntimed = kernel_window(sim, min(args.steps, 1000))
            pairs = sim.count_pairs(2.0 * h) if prec == "fp32" else None
            lay = sim.layout()
            st = sim.read()
            state = (st["position"].double().cpu().numpy(), st["velocity"].double().cpu().numpy())
        finally:
            sim.close()
        runs[prec] = {"wall_s": wall_s, "rebuilds": rebuilds, "sweep": sweep, "timed_launches": ntimed,
                      "pairs": pairs, "state": state, "warm": warm, "t_start": t_start, "layout": lay}
        print(f"dam break {prec}: {1000 * wall_s / args.steps:.4f} ms/step, {rebuilds} rebuilds in "
              f"{args.steps} steps", file=sys.stderr, flush=True)
    r32 = runs["fp32"]
    value = nf * args.steps / r32["wall_s"]
    ms_step = 1000.0 * r32["wall_s"] / args.steps
    # algorithmic bytes per launch of the force sweep: reads x, v (FP64, 2 x 16 B),
    # m, rho (4 B each), writes x, v (2 x 16 B) per fluid particle, plus each wall
    # particle's x, m, rho, p read once (16 + 3 x 4 B)
    s = 4
    b_force = nf * (32 * 2 + 2 * s) + nw * (16 + 3 * s)
    force_ms = r32["sweep"][2]
    achieved = b_force / (force_ms * 1e-3) / 1e9 if force_ms > 0 else 0.0
    # step bytes of the compulsory two-sweep model: 40 D + 4 s per fluid particle,
    # 24 D + 2 s per wall particle
    step_bytes = nf * (40 * 2 + 4 * s) + nw * (24 * 2 + 2 * s)
    on_halo = tuned["best"] == "halo"
    kernel = "k_step_force" if tuned["best"] in ("lists", "halo") else "k_tile_force"
    prof = PROFILE_HALO if on_halo else (PROFILE_LISTS if kernel == "k_step_force" else PROFILE)
    traffic = None if args.diffusion else profiled_traffic(nf, prof, kernel)
    result = {
        "metric": "particle-updates/s (2-D WCSPH dam break, FP32-compensated) + HBM GB/s",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": r32["warm"],
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: deterministic dam-break lattice (no dataset)",
        "config": {"workload": f"2-D dam break, {nf} fluid + {nw} wall particles, "
                               f"dx={cfg.dx:.5g} m, h=1.3dx, c0={cfg.sound_speed:.4g} m/s, "
                               f"dt={cfg.dt:.4g} s, nu={cfg.viscosity}, skin 0.25h"
                               + (f", Molteni δ-diffusion δ={args.diffusion}" if args.diffusion else ""),
                   "precision": "pair arithmetic FP32 with Kahan sums; positions, velocities "
                                "and integration FP64",
                   "fluid": nf, "wall": nw, "parallelism": "single GPU"},
        "window": {"start_time_s": r32["t_start"], "untimed_steps_before": r32["warm"],
                   "list_rebuilds": r32["rebuilds"],
                   "steps_per_rebuild": args.steps / r32["rebuilds"] if r32["rebuilds"] else None,
                   "how": "the collapse developed untimed to start_time_s; the window's wall clock "
                          "includes every list rebuild in it"},
        "pair_interactions_per_s": r32["pairs"] * args.steps / r32["wall_s"],
        "pairs_within_2h": r32["pairs"],
        "layout": {"chosen": tuned["best"], "trial_ms_per_step": tuned["ms"],
                   "halo_tiles": r32["layout"]["halo_tiles"], "tiles": r32["layout"]["tiles"],
                   "largest_halo_rows": r32["layout"]["halo_max"]},
        "algorithmic_GBps_step": step_bytes * args.steps / r32["wall_s"] / 1e9,
        "algorithmic_bytes_per_step": step_bytes,
        "step_roofline_frac": step_bytes * args.steps / r32["wall_s"] / 1e9 / HBM_PEAK_GBS,
        "sweep_ms_per_launch": {"density": r32["sweep"][0], "wall": r32["sweep"][1],
                                "force_integrate": force_ms, "timed_launches": r32["timed_launches"],
                                "how": "HIP events around each sweep on every 10th step of a "
                                       "second window on the stepping stream"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": None if traffic is None else os.path.relpath(prof, HERE),
                     "kernel": f"{kernel} ({tuned['best']} layout: force + viscosity + integrate)",
                     "algorithmic_bytes_per_launch": b_force},
    }
    if "fp64" in runs:
        r64 = runs["fp64"]
        v64 = nf * args.steps / r64["wall_s"]
        b64 = nf * (32 * 2 + 2 * 8) + nw * (16 + 3 * 8)
        f64_ms = r64["sweep"][2]
        result["fp64"] = {"value": v64, "ms_per_step": 1000.0 * r64["wall_s"] / args.steps,
                          "list_rebuilds": r64["rebuilds"], "start_time_s": r64["t_start"],
                          "sweep_ms_per_launch": {"density": r64["sweep"][0], "wall": r64["sweep"][1],
                                                  "force_integrate": f64_ms},
                          "force_sweep_GBps": b64 / (f64_ms * 1e-3) / 1e9 if f64_ms > 0 else 0.0}
        result["fp32c_over_fp64"] = value / v64
        result["fp32c_vs_fp64_error"] = precision_error(cfg, mass, wall, wmass, r64["state"], args.diffusion)
    if not args.no_serve:
        result["serving"] = serving_bench(args.serve_requests, args.serve_steps, args.serve_sequential,
                                          args.serve_sequential_api)
    if not args.no_3d:
        blk = argparse.Namespace(**vars(args))
        blk.steps, blk.warmup = args.block_steps, args.block_warmup
        blk.decay_n, blk.t_end = 0, 0.0
        result["shear_3d"] = shear_bench(blk)
        result["box_3d"] = kleefsman_bench(blk)
    if not args.no_cpu:
        st = runs["fp64"]["state"] if "fp64" in runs else runs["fp32"]["state"]
        result["cpu_baseline"] = cpu_baseline(cfg, wmass, wall, (st[0], st[1], mass), args.cpu_seconds)
    return result


if __name__ == "__main__":
    main()
