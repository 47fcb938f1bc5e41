"""3-D dam break against a box obstacle (kleefsman_obstacle_3d.py:553-726) on
the resident stepper.

``kleefsman_state`` builds the reference's particle sets (column, tank walls
and the box, kleefsman_obstacle_3d.py:570-577); ``run_case`` advances them
with rhs_3d physics, the Verlet pair cache, symplectic Euler,
apply_boundary_3d and clamp_out_of_obstacle, all resident on the GPU, and
records what the reference's driver records: the water-height gauges (slab,
peak height, centreline H2, front, bypass count), the pressure taps on the box
(point or disk cubature), the force, moment and impulses on the box, and
crash-safe checkpoints (kleefsman_obstacle_3d.py:273-726,
marin_probes.py:1-83).  ``run_loads`` is the loads-only loop.  The gauges
and taps are reductions on the device (tsph_gauges / tsph_sim_gauges).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .cases import make_column_3d, mass_3d, obstacle_lattice_3d, wall_lattice_3d
from .config import Kleefsman3DConfig
from .engine import METER_KEYS, WcsphSimulation

# SPHERIC Test 2 workbook, H2 (marin_probes.py:7-12)
H2_PEAK_M = .3408
H2_PEAK_TIME_S = 1.7870422657819276
H2_ARRIVAL_S = .3673968228757868


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


def make_simulation(cfg: Kleefsman3DConfig, precision: str = "fp32", device="cuda", pos=None,
                    vel=None, adaptive=None):
    """The resident stepper of the case; ``pos`` / ``vel`` (fluid, caller
    order) replace the initial column, as a resumed run needs."""
    pos0, mass, wall, rows, centre, box = kleefsman_state(cfg)
    pos = pos0 if pos is None else np.asarray(pos, np.float64)
    vel = np.zeros_like(pos) if vel is None else np.asarray(vel, np.float64)
    if pos.shape != pos0.shape or vel.shape != pos0.shape:
        raise ValueError(f"state holds {pos.shape} particles, this configuration builds {pos0.shape}")
    h = cfg.h_factor * cfg.dx
    wall_mass = mass_3d(cfg.dx, cfg.h_factor, cfg.rho0, cfg.density_calibration)
    if cfg.density_diffusion > 0.0 and cfg.density_diffusion_form != "molteni":
        # rhs_3d takes the Molteni form only (dam_break_3d.py:330-336)
        raise ValueError("density_diffusion_rate implements the 'molteni' form only; the "
                         f"configured form is {cfg.density_diffusion_form!r}")
    sim = WcsphSimulation(pos, vel, mass, wall, h=h, rho0=cfg.rho0,
                          sound_speed=cfg.sound_speed, gamma=cfg.gamma, gravity=cfg.gravity,
                          viscosity=cfg.viscosity, dt=cfg.dt, wall_mass=wall_mass,
                          skin=cfg.pair_skin_factor * h, lo=[0.0, 0.0, 0.0],
                          hi=[cfg.pool_width, cfg.tank_top(), cfg.tank_depth], tol=0.25 * cfg.dx,
                          wall_friction=cfg.wall_friction, density_diffusion=cfg.density_diffusion,
                          diffusion_form="molteni", precision=precision, device=device,
                          obstacle={"rows": rows, "centre": centre, "box": box, "eps": 0.01 * cfg.dx},
                          adaptive=adaptive)
    return sim, pos.shape[0], wall.shape[0]


def run_loads(cfg: Kleefsman3DConfig, steps: int | None = None, precision: str = "fp32",
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


# --------------------------------------------------------------------------- instruments

def sensor_positions(cfg: Kleefsman3DConfig) -> np.ndarray:
    """The eight pressure sensors (marin_probes.py:15-20): four on the
    upstream face, four on the top face, in solver coordinates."""
    offsets = [0.021 + 0.04 * i for i in range(4)]
    front = [[cfg.obstacle_x, s, cfg.obstacle_z + 0.176] for s in offsets]
    top = [[cfg.obstacle_x + s, cfg.obstacle_height, cfg.obstacle_z1() - 0.176] for s in offsets]
    return np.array(front + top, np.float64)


def tap_disk_nodes(cfg: Kleefsman3DConfig, taps, diameter: float) -> np.ndarray:
    """Centre and six azimuthal nodes of the fifth-degree equal-area disk
    cubature on each sensor's face, (8, 7, 3); (8, 1, 3) for diameter 0
    (marin_probes.py:30-48)."""
    taps = np.asarray(taps, np.float64)
    if not math.isfinite(diameter) or diameter < 0:
        raise ValueError("pressure disk diameter must be finite and nonnegative")
    if diameter == 0:
        return taps[:, None, :].copy()
    radius = diameter / 2
    clearance = np.concatenate((taps[:4, 1], cfg.obstacle_height - taps[:4, 1],
                                taps[4:, 0] - cfg.obstacle_x, cfg.obstacle_x1() - taps[4:, 0],
                                taps[:, 2] - cfg.obstacle_z, cfg.obstacle_z1() - taps[:, 2]))
    if radius > float(clearance.min()) + 1e-7:
        raise ValueError("pressure disk extends beyond the instrumented face")
    angle = np.arange(6) * (math.pi / 3)
    ring = radius * math.sqrt(2 / 3)
    nodes = np.repeat(taps[:, None, :], 7, axis=1)
    nodes[:4, 1:, 1] += ring * np.cos(angle)
    nodes[4:, 1:, 0] += ring * np.cos(angle)
    nodes[:, 1:, 2] += ring * np.sin(angle)
    return nodes


def disk_average(node_pressures) -> np.ndarray:
    """Disk pressure from the seven node pressures (marin_probes.py:72-74)."""
    p = np.asarray(node_pressures, np.float64)
    return .25 * p[:, 0] + .125 * p[:, 1:].sum(1)


def gauge_spec(cfg: Kleefsman3DConfig, nodes) -> _lib.GaugeSpec:
    nodes = np.asarray(nodes, np.float64)
    if nodes.ndim != 3 or nodes.shape[0] != 8 or nodes.shape[1] not in (1, 7) or nodes.shape[2] != 3:
        raise ValueError(f"nodes must be (8, 1 or 7, 3), got {nodes.shape}")
    s = _lib.GaugeSpec()
    s.station_x, s.aperture = cfg.station_x(), cfg.gauge_aperture
    s.tank_depth, s.dx, s.h, s.gravity = cfg.tank_depth, cfg.dx, cfg.h_factor * cfg.dx, cfg.gravity
    for d, (a, b) in enumerate(((cfg.obstacle_x, cfg.obstacle_x1()), (0.0, cfg.obstacle_height),
                                (cfg.obstacle_z, cfg.obstacle_z1()))):
        s.obstacle_lo[d], s.obstacle_hi[d] = a, b
    s.per_tap = nodes.shape[1]
    flat = nodes.reshape(-1)
    for k in range(flat.size):
        s.nodes[k] = float(flat[k])
    return s


def _device_gauges(cfg, pos, volume=None, pressure=None, density=None, nodes=None) -> np.ndarray:
    """tsph_gauges on caller tensors (on a ROCm device)."""
    dev = _lib.require_gpu(pos)
    n = pos.shape[0]
    sdt = volume.dtype if volume is not None else torch.float64
    code = _lib.F32 if sdt == torch.float32 else _lib.F64
    x = pos.detach().to(torch.float64).contiguous()
    fill = lambda v, c: (torch.full((n,), c, dtype=sdt, device=dev) if v is None
                         else v.detach().to(sdt).contiguous())
    vol, rho, p = fill(volume, 0.0), fill(density, 1.0), fill(pressure, 0.0)
    spec = gauge_spec(cfg, sensor_positions(cfg)[:, None, :] if nodes is None else nodes)
    out = np.zeros(_lib.GAUGE_ROW, np.float64)
    ctx = _lib.default_context(dev)
    _lib.check(_lib.lib().tsph_gauges(ctx.handle, C.byref(spec), code, n, _lib.ptr(x), _lib.ptr(vol),
                                      _lib.ptr(rho), _lib.ptr(p),
                                      out.ctypes.data_as(C.POINTER(C.c_double)),
                                      _lib.stream_handle(dev)))
    return out


def slab_height(pos: torch.Tensor, cfg: Kleefsman3DConfig) -> float:
    """n dx^3 / (aperture * obstacle width) over the gauge slab
    (kleefsman_obstacle_3d.py:281-295)."""
    return float(_device_gauges(cfg, pos)[0])


def peak_height(pos: torch.Tensor, cfg: Kleefsman3DConfig) -> float:
    """Highest wetted point in the slab, 0 if it is dry (:298-306)."""
    return float(_device_gauges(cfg, pos)[1])


def bypass_count(pos: torch.Tensor, cfg: Kleefsman3DConfig) -> int:
    """Particles past the box outside its spanwise extent (:309-318)."""
    return int(_device_gauges(cfg, pos)[3])


def centreline_height(pos: torch.Tensor, volume: torch.Tensor, cfg: Kleefsman3DConfig) -> float:
    """Water volume per horizontal area under a 2-D kernel centred on the H2
    station (marin_probes.py:23-27)."""
    return float(_device_gauges(cfg, pos, volume=volume)[4])


def tap_node_pressures(pos, pressure, density, volume, cfg: Kleefsman3DConfig, nodes):
    """One-sided SPH pressure at each quadrature node, (8, K) each for the
    pressure and the support count (marin_probes.py:51-69)."""
    nodes = np.asarray(nodes, np.float64)
    row = _device_gauges(cfg, pos, volume=volume, pressure=pressure, density=density, nodes=nodes)
    k = nodes.shape[1]
    pv = row[5:5 + 2 * 8 * k:2].reshape(8, k)
    sv = row[6:6 + 2 * 8 * k:2].reshape(8, k).astype(np.int64)
    return pv, sv


def tap_pressures(pos, pressure, density, volume, cfg: Kleefsman3DConfig, taps=None):
    """Point pressure and support at the eight sensor centres (marin_probes.py:77-80)."""
    taps = sensor_positions(cfg) if taps is None else np.asarray(taps, np.float64)
    p, count = tap_node_pressures(pos, pressure, density, volume, cfg, taps[:, None, :])
    return p[:, 0], count[:, 0]


def arrival_time(times, values, threshold: float) -> float | None:
    """First upward crossing of ``threshold``, linearly interpolated (:325-334)."""
    for k in range(1, len(values)):
        if values[k - 1] < threshold <= values[k]:
            t0, t1 = times[k - 1], times[k]
            v0, v1 = values[k - 1], values[k]
            if v1 == v0:
                return float(t1)
            return float(t0 + (threshold - v0) * (t1 - t0) / (v1 - v0))
    return None


def _lattice_extent(extent: float, dx: float) -> float:
    return max(1, int(round(extent / dx))) * dx


# --------------------------------------------------------------------------- run record

@dataclass
class Kleefsman3DRun:
    """The record of one run (kleefsman_obstacle_3d.py:337-458)."""
    cfg: Kleefsman3DConfig
    times: list = field(default_factory=list)
    slab: list = field(default_factory=list)
    local_h2: list = field(default_factory=list)
    height: list = field(default_factory=list)
    front: list = field(default_factory=list)
    bypass: list = field(default_factory=list)
    sph_impulse: list = field(default_factory=lambda: [0., 0., 0.])
    clamp_impulse: list = field(default_factory=lambda: [0., 0., 0.])
    max_speed: float = 0.0
    max_density_compression: float = 0.0
    boundary_escapes: int = 0
    obstacle_clamps: int = 0
    mass_error_rel: float = 0.0
    fluid_particles: int = 0
    wall_particles: int = 0
    steps: int = 0
    simulated_time: float = 0.0
    wall_clock_s: float = 0.0
    resumed: bool = False

    def peak(self):
        if not self.slab:
            return 0.0, 0.0
        k = max(range(len(self.slab)), key=lambda i: self.slab[i])
        return float(self.slab[k]), float(self.times[k])

    def arrival(self) -> float | None:
        return arrival_time(self.times, self.slab, self.cfg.arrival_threshold)

    def operator_consistency(self) -> dict:
        """Samples with more than 2 dx of water in the slab but a highest wetted
        point of 0 cannot come from one particle set (:380-412)."""
        zero_wet = [s for s, h in zip(self.slab, self.height) if s > 2.0 * self.cfg.dx and h <= 0.0]
        wet = sum(1 for s in self.slab if s > 0.0)
        return {"wet_samples": int(wet), "impossible_samples": len(zero_wet),
                "worst_impossible_slab_m": max(zero_wet) if zero_wet else 0.0,
                "height_column_admissible": not zero_wet,
                "reason": ("" if not zero_wet else
                           "samples report slab > 2*dx of water while the highest wetted point "
                           "in the same band is 0.0")}

    def to_dict(self) -> dict:
        cfg = self.cfg
        peak_h, peak_t = self.peak()
        return {
            "dx": cfg.dx, "t_end": cfg.t_end, "dt": cfg.dt, "sound_speed": cfg.sound_speed,
            "fluid_particles": self.fluid_particles, "wall_particles": self.wall_particles,
            "blocked_fraction": cfg.blocked_fraction(), "obstacle_z": cfg.obstacle_z,
            "obstacle_width_nominal_m": cfg.obstacle_width,
            "obstacle_width_on_lattice_m": _lattice_extent(cfg.obstacle_width, cfg.dx),
            "stations_x_m": [cfg.station_x()], "gauge_aperture_m": cfg.gauge_aperture,
            "peak_height_m": peak_h,
            "local_h2_peak_m": max(self.local_h2) if self.local_h2 else None,
            "local_h2_arrival_s": (arrival_time(self.times, self.local_h2, cfg.arrival_threshold)
                                   if self.local_h2 else None),
            "peak_time_s": peak_t,
            "peak_height_operator_m": max(self.height) if self.height else 0.0,
            "arrival_time_s": self.arrival(), "arrival_threshold_m": cfg.arrival_threshold,
            "max_speed": self.max_speed, "max_density_compression": self.max_density_compression,
            "boundary_escapes": self.boundary_escapes, "obstacle_clamps": self.obstacle_clamps,
            "obstacle_sph_impulse_Ns": self.sph_impulse,
            "obstacle_clamp_impulse_Ns": self.clamp_impulse,
            "mass_error_rel": self.mass_error_rel, "steps": self.steps,
            "simulated_time_s": self.simulated_time, "wall_clock_s": self.wall_clock_s,
            "max_bypass_particles": max(self.bypass) if self.bypass else 0.0,
            "fingerprint": {
                "t_end": cfg.t_end, "sound_speed": cfg.sound_speed, "wall_layers": cfg.wall_layers,
                "wall_operator": "finite-volume-self-v2", "viscosity": cfg.viscosity,
                "obstacle_x": cfg.obstacle_x, "obstacle_z": cfg.obstacle_z,
                "station_offset": cfg.station_offset, "gauge_aperture": cfg.gauge_aperture,
                "pressure_disk_diameter": cfg.pressure_disk_diameter,
                "arrival_threshold": cfg.arrival_threshold,
            },
        }


# --------------------------------------------------------------------------- checkpoints

CHECKPOINT_VERSION = 3

CHECKPOINT_FINGERPRINT_FIELDS = (
    "dx", "sound_speed", "gamma", "rho0", "h_factor", "wall_layers",
    "obstacle_z", "obstacle_width", "station_offset", "gauge_aperture",
    "arrival_threshold", "pair_skin_factor", "density_diffusion",
    "surface_pressure_closure", "adaptive_dt", "max_courant", "viscosity",
    "pool_width", "tank_depth", "tank_height", "column_width",
    "column_height", "column_depth", "obstacle_x", "obstacle_length",
    "obstacle_height", "gravity", "wall_friction", "dt", "output_every",
)


def checkpoint_fingerprint(cfg) -> dict:
    """The configuration fields that decide whether a checkpoint is reusable
    (kleefsman_obstacle_3d.py:484-490); t_end is a stopping rule, not physics."""
    fields = {name: getattr(cfg, name) for name in CHECKPOINT_FINGERPRINT_FIELDS}
    if cfg.pressure_disk_diameter:
        fields["pressure_disk_diameter"] = cfg.pressure_disk_diameter
    return fields


def save_checkpoint(path, cfg, *, step, elapsed, dt, positions, velocity, run, speeds, compression,
                    escapes, clamps) -> None:
    """Write the integration state atomically: a temporary file, then
    os.replace (:493-517).  Same payload keys and version as the reference."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    payload = {
        "version": CHECKPOINT_VERSION, "fingerprint": checkpoint_fingerprint(cfg),
        "step": int(step), "elapsed": float(elapsed), "dt": float(dt),
        "positions": torch.as_tensor(positions).detach().to("cpu").clone(),
        "velocity": torch.as_tensor(velocity).detach().to("cpu").clone(),
        "times": list(run.times), "height": list(run.height), "local_h2": list(run.local_h2),
        "front": list(run.front), "bypass": list(run.bypass), "speeds": float(speeds),
        "compression": float(compression), "escapes": int(escapes), "clamps": int(clamps),
        "sph_impulse": list(run.sph_impulse), "clamp_impulse": list(run.clamp_impulse),
        "slab": list(run.slab),
    }
    tmp = path.with_name(path.name + ".tmp")
    torch.save(payload, tmp)
    os.replace(tmp, path)


def load_checkpoint(path, cfg) -> dict:
    """Load a checkpoint, refusing one written for different physics (:520-535).
    The file is read with torch.load(weights_only=True): tensors, lists,
    dicts and numbers only."""
    payload = torch.load(path, map_location="cpu", weights_only=True)
    if payload.get("version") != CHECKPOINT_VERSION:
        raise ValueError(f"checkpoint version {payload.get('version')} != {CHECKPOINT_VERSION}")
    got = payload.get("fingerprint", {})
    want = checkpoint_fingerprint(cfg)
    diff = {k: (got.get(k), want.get(k)) for k in got.keys() | want.keys() if got.get(k) != want.get(k)}
    if diff:
        raise ValueError("checkpoint was written for a different configuration: "
                         + ", ".join(f"{k}: {g!r} -> {w!r}" for k, (g, w) in diff.items()))
    return payload


# --------------------------------------------------------------------------- driver

def run_case(cfg: Kleefsman3DConfig, device="cuda", progress=None, on_sample=None,
             checkpoint_path=None, checkpoint_every: int = 0, resume: bool = True,
             precision: str = "fp64") -> Kleefsman3DRun:
    """Integrate to cfg.t_end and record the gauge series
    (kleefsman_obstacle_3d.py:553-726).

    The reference's loop evaluates step k at the state it begins with,
    samples that state when k is a multiple of ``output_every`` or the final
    step (simulated time >= t_end), then integrates; a checkpoint every
    ``checkpoint_every`` steps holds the state after the integration.  The
    stepper here runs whole steps, so a sample is read from the input side of
    the step just run (tsph_sim_gauges) and a checkpoint from its output side;
    the loads, clamps and escapes are those of the steps the reference
    integrates.  ``on_sample(i, sample)`` gets the reference's sample dict.
    ``precision``: "fp64" or "fp32" (FP32-compensated sweeps)."""
    if str(device) == "cpu":
        device = "cuda"
    if getattr(cfg, "adaptive_dt", False):
        return _run_case_adaptive(cfg, device, progress, on_sample, precision, checkpoint_path,
                                  checkpoint_every, resume)
    t_start = time.time()
    pos0, mass, wall, rows, centre, box = kleefsman_state(cfg)
    nf, nw = pos0.shape[0], wall.shape[0]
    run = Kleefsman3DRun(cfg=cfg, fluid_particles=nf, wall_particles=nw)
    dt, elapsed, start_step = cfg.dt, 0.0, 0
    max_steps = cfg.max_steps or cfg.steps_for_t_end()
    speeds = compression = 0.0
    base_escapes = base_clamps = 0
    base_sph, base_clamp = np.zeros(3), np.zeros(3)
    positions, velocity = pos0, np.zeros_like(pos0)
    resumed = False
    if checkpoint_path is not None and resume and Path(checkpoint_path).exists():
        state = load_checkpoint(checkpoint_path, cfg)
        if tuple(state["positions"].shape) != pos0.shape:
            raise ValueError(f"checkpoint holds {tuple(state['positions'].shape)} particles, "
                             f"this configuration builds {pos0.shape}")
        positions = state["positions"].double().numpy()
        velocity = state["velocity"].double().numpy()
        dt, elapsed, start_step = float(state["dt"]), float(state["elapsed"]), int(state["step"]) + 1
        run.times, run.height = list(state["times"]), list(state["height"])
        run.local_h2, run.front = list(state["local_h2"]), list(state["front"])
        run.bypass, run.slab = list(state["bypass"]), list(state.get("slab", []))
        speeds, compression = float(state["speeds"]), float(state["compression"])
        base_escapes, base_clamps = int(state["escapes"]), int(state["clamps"])
        base_sph = np.array(state["sph_impulse"], np.float64)
        base_clamp = np.array(state["clamp_impulse"], np.float64)
        resumed = True
        if progress is not None:
            progress(elapsed, run.slab[-1] if run.slab else 0.0)
    # the reference's clock: time at the start of each step, the final step
    times, t, final_step = {}, elapsed, None
    for step in range(start_step, max_steps + 1):
        times[step] = t
        if t >= cfg.t_end - 1.0e-12:
            final_step = step
            break
        t += dt
    last_int = max_steps if final_step is None else final_step - 1      # last integrated step
    last_eval = max_steps if final_step is None else final_step
    events = [(s, 0) for s in range(start_step, last_eval + 1)
              if s % cfg.output_every == 0 or s == final_step]
    if checkpoint_path is not None and checkpoint_every > 0:
        events += [(c, 1) for c in range(start_step, last_int + 1) if c % checkpoint_every == 0]
    events.append((last_int, 2))
    events.sort()
    nodes = tap_disk_nodes(cfg, sensor_positions(cfg), cfg.pressure_disk_diameter)
    spec = gauge_spec(cfg, nodes)
    kn = nodes.shape[1]
    sim, _, _ = make_simulation(cfg, precision, device, pos=positions, vel=velocity)
    done = 0                          # steps run: start_step .. start_step + done - 1

    def advance(through: int) -> None:
        nonlocal done
        want = through + 1 - start_step
        if want > done:
            sim.step(want - done)
            done = want

    def totals(through: int):
        """sph impulse, clamp impulse and clamp count after integrating steps
        start_step .. through."""
        lr = sim.loads()[:max(through + 1 - start_step, 0)]
        if not len(lr):
            return base_sph.copy(), base_clamp.copy(), base_clamps
        return (base_sph + (lr[:, 1:4] * dt).sum(0), base_clamp + lr[-1, 7:10],
                base_clamps + int(lr[-1, 10]))

    escapes = base_escapes
    try:
        for step, kind in events:
            if kind == 2:
                if last_int >= start_step:
                    advance(last_int)
                escapes = base_escapes + sim.stats()["escaped"]
                continue
            advance(step)
            if kind == 1:
                st = sim.stats()
                sph, clamp_imp, clamps = totals(step)
                cur = sim.read()
                ck = Kleefsman3DRun(cfg=cfg)
                ck.times, ck.slab, ck.height = run.times, run.slab, run.height
                ck.local_h2, ck.front, ck.bypass = run.local_h2, run.front, run.bypass
                ck.sph_impulse, ck.clamp_impulse = sph.tolist(), clamp_imp.tolist()
                save_checkpoint(checkpoint_path, cfg, step=step, elapsed=times[step] + dt, dt=dt,
                                positions=cur["position"], velocity=cur["velocity"], run=ck,
                                speeds=max(speeds, st["peak_speed"]),
                                compression=max(compression, st["peak_rho"] / cfg.rho0 - 1.0),
                                escapes=base_escapes + st["escaped"], clamps=clamps)
                continue
            g = sim.gauges(spec)
            t_s = times[step]
            run.times.append(t_s)
            run.slab.append(float(g[0]))
            run.height.append(float(g[1]))
            run.front.append(float(g[2]))
            run.bypass.append(float(g[3]))
            run.local_h2.append(float(g[4]))
            if on_sample is not None:
                lr = sim.loads()[step - start_step]
                sph, clamp_imp, _ = totals(step - 1)
                pv = g[5:5 + 16 * kn:2].reshape(8, kn)
                sv = g[6:6 + 16 * kn:2].reshape(8, kn).astype(np.int64)
                on_sample(len(run.times) - 1, {
                    "obstacle_sph_force_N": lr[1:4].tolist(),
                    "obstacle_sph_impulse_Ns": sph.tolist(),
                    "obstacle_clamp_impulse_Ns": clamp_imp.tolist(),
                    "H2_local_m": run.local_h2[-1],
                    "tap_pressure_Pa": pv[:, 0].tolist(), "tap_support": sv[:, 0].tolist(),
                    "disk_pressure_Pa": disk_average(pv).tolist() if cfg.pressure_disk_diameter else [],
                    "disk_support": sv.min(1).tolist() if cfg.pressure_disk_diameter else [],
                    "obstacle_sph_moment_Nm": lr[4:7].tolist(),
                    "t_s": t_s, "slab_height_m": run.slab[-1], "height_operator_m": run.height[-1],
                    "front_x_m": run.front[-1], "n_bypass": run.bypass[-1],
                })
            if progress is not None and step % (cfg.output_every * 20) == 0:
                progress(t_s, run.slab[-1])
        advance(last_eval)
        sph, clamp_imp, clamps = totals(last_int)
        st = sim.stats()
    finally:
        sim.close()
    run.sph_impulse, run.clamp_impulse = sph.tolist(), clamp_imp.tolist()
    run.max_speed = max(speeds, st["peak_speed"])
    run.max_density_compression = max(compression, st["peak_rho"] / cfg.rho0 - 1.0)
    run.boundary_escapes = escapes
    run.obstacle_clamps = clamps
    run.mass_error_rel = 0.0            # the particle masses are constant
    run.steps = last_eval
    run.simulated_time = times[final_step] if final_step is not None else t
    run.wall_clock_s = time.time() - t_start
    run.resumed = resumed
    if checkpoint_path is not None:
        try:
            os.unlink(checkpoint_path)
        except FileNotFoundError:
            pass
    return run


def _sequential_impulse(base, loads) -> np.ndarray:
    """base + sum_k F_k dt_k over load rows, added one step at a time in step
    order, as the reference accumulates its impulse tensor."""
    acc = np.array(base, np.float64)
    for row in loads:
        acc = acc + row[1:4] * row[11]
    return acc


def _run_case_adaptive(cfg: Kleefsman3DConfig, device, progress, on_sample, precision,
                       checkpoint_path=None, checkpoint_every: int = 0, resume: bool = True):
    """run_case with adaptive_dt (kleefsman_obstacle_3d.py:644-713): after each
    step dt = next_dt(dt, largest speed at its start), the step is never
    shortened onto t_end, and the first step starting at elapsed >= t_end -
    1e-12 is the final sample; the loop also ends after max_steps.

    Checkpoints (:597-628, :701-706) hold the state after each step s with
    s % checkpoint_every == 0, the simulated time at the start of step s + 1
    and the step-size variable the rule carries into it; a resumed run starts
    the stepper from that time state (tsph_sim_params.initial_time /
    initial_dt) and numbers its steps from s + 1."""
    from .config import adaptive_controls
    t_start = time.time()
    pos0, mass, wall, rows_, centre, box = kleefsman_state(cfg)
    run = Kleefsman3DRun(cfg=cfg, fluid_particles=pos0.shape[0], wall_particles=wall.shape[0])
    max_steps = cfg.max_steps or cfg.steps_for_t_end()
    ctl = dict(adaptive_controls(cfg), target_time=float(cfg.t_end), free_last=True, tol=1.0e-12)
    positions, velocity = pos0, np.zeros_like(pos0)
    start_step, t0 = 0, 0.0
    speeds = compression = 0.0
    base_escapes = base_clamps = 0
    base_sph, base_clamp = np.zeros(3), np.zeros(3)
    resumed = False
    if checkpoint_path is not None and resume and Path(checkpoint_path).exists():
        state = load_checkpoint(checkpoint_path, cfg)
        if tuple(state["positions"].shape) != pos0.shape:
            raise ValueError(f"checkpoint holds {tuple(state['positions'].shape)} particles, "
                             f"this configuration builds {pos0.shape}")
        positions = state["positions"].double().numpy()
        velocity = state["velocity"].double().numpy()
        t0 = float(state["elapsed"])
        ctl["initial_time"], ctl["initial_dt"] = t0, float(state["dt"])
        start_step = int(state["step"]) + 1
        run.times, run.height = list(state["times"]), list(state["height"])
        run.local_h2, run.front = list(state["local_h2"]), list(state["front"])
        run.bypass, run.slab = list(state["bypass"]), list(state.get("slab", []))
        speeds, compression = float(state["speeds"]), float(state["compression"])
        base_escapes, base_clamps = int(state["escapes"]), int(state["clamps"])
        base_sph = np.array(state["sph_impulse"], np.float64)
        base_clamp = np.array(state["clamp_impulse"], np.float64)
        resumed = True
        if progress is not None:
            progress(t0, run.slab[-1] if run.slab else 0.0)
    nodes = tap_disk_nodes(cfg, sensor_positions(cfg), cfg.pressure_disk_diameter)
    spec = gauge_spec(cfg, nodes)
    kn = nodes.shape[1]
    ow = cfg.output_every
    ce = int(checkpoint_every) if checkpoint_path is not None else 0
    sim, _, _ = make_simulation(cfg, precision, device, pos=positions, vel=velocity, adaptive=ctl)
    clock = [0, t0]                       # [local step, simulated time at its start]

    def time_at(local, lr):
        """Simulated time at the start of local step ``local``: the stepper's
        sum of step sizes, in the same order."""
        k, t = clock
        if local < k:
            k, t = 0, t0
        for j in range(k, local):
            t += float(lr[j, 11])
        clock[0], clock[1] = local, t
        return t

    def totals(lr):
        """(sph impulse, clamp impulse, clamp count) after the load rows lr."""
        if not len(lr):
            return base_sph.copy(), base_clamp.copy(), base_clamps
        return (_sequential_impulse(base_sph, lr), base_clamp + lr[-1, 7:10],
                base_clamps + int(lr[-1, 10]))

    try:
        while True:
            st = sim.stats()
            nxt = start_step + st["steps"]          # the next step to run (global)
            if st["finished"] or nxt > max_steps:
                break
            through = ((nxt + ow - 1) // ow) * ow
            if ce > 0:
                through = min(through, ((nxt + ce - 1) // ce) * ce)
            through = min(through, max_steps)
            sim.step(through + 1 - nxt, meter_every=0)
            st = sim.stats()
            local = st["steps"] - 1               # the step just run
            step = start_step + local
            if st["finished"] or step % ow == 0:
                g = sim.gauges(spec)
                lr = sim.loads()
                t_s = time_at(local, lr)
                run.times.append(t_s)
                run.slab.append(float(g[0]))
                run.height.append(float(g[1]))
                run.front.append(float(g[2]))
                run.bypass.append(float(g[3]))
                run.local_h2.append(float(g[4]))
                if on_sample is not None:
                    sph, clamp_imp, _ = totals(lr[:local])
                    pv = g[5:5 + 16 * kn:2].reshape(8, kn)
                    sv = g[6:6 + 16 * kn:2].reshape(8, kn).astype(np.int64)
                    on_sample(len(run.times) - 1, {
                        "obstacle_sph_force_N": lr[local, 1:4].tolist(),
                        "obstacle_sph_impulse_Ns": sph.tolist(),
                        "obstacle_clamp_impulse_Ns": clamp_imp.tolist(),
                        "H2_local_m": run.local_h2[-1],
                        "tap_pressure_Pa": pv[:, 0].tolist(), "tap_support": sv[:, 0].tolist(),
                        "disk_pressure_Pa": disk_average(pv).tolist() if cfg.pressure_disk_diameter else [],
                        "disk_support": sv.min(1).tolist() if cfg.pressure_disk_diameter else [],
                        "obstacle_sph_moment_Nm": lr[local, 4:7].tolist(),
                        "t_s": t_s, "slab_height_m": run.slab[-1], "height_operator_m": run.height[-1],
                        "front_x_m": run.front[-1], "n_bypass": run.bypass[-1],
                    })
                if progress is not None and step % (ow * 20) == 0:
                    progress(t_s, run.slab[-1])
            if ce > 0 and step % ce == 0 and not st["finished"]:
                ts = sim.time_state()
                cur = sim.read()
                sph, clamp_imp, clamps = totals(sim.loads())
                ck = Kleefsman3DRun(cfg=cfg)
                ck.times, ck.slab, ck.height = run.times, run.slab, run.height
                ck.local_h2, ck.front, ck.bypass = run.local_h2, run.front, run.bypass
                ck.sph_impulse, ck.clamp_impulse = sph.tolist(), clamp_imp.tolist()
                save_checkpoint(checkpoint_path, cfg, step=step, elapsed=ts["elapsed"], dt=ts["dt"],
                                positions=cur["position"], velocity=cur["velocity"], run=ck,
                                speeds=max(speeds, st["peak_speed"]),
                                compression=max(compression, st["peak_rho"] / cfg.rho0 - 1.0),
                                escapes=base_escapes + st["escaped"], clamps=clamps)
        st = sim.stats()
        lr = sim.loads()
    finally:
        sim.close()
    final = bool(st["finished"])
    last_eval = st["steps"] - 1
    integ = lr[:last_eval] if final else lr          # the final sample is evaluated, not moved
    sph, clamp_imp, clamps = totals(integ)
    run.sph_impulse, run.clamp_impulse = sph.tolist(), clamp_imp.tolist()
    run.obstacle_clamps = clamps
    run.max_speed = max(speeds, st["peak_speed"])
    run.max_density_compression = max(compression, st["peak_rho"] / cfg.rho0 - 1.0)
    run.boundary_escapes = base_escapes + st["escaped"]
    run.steps = start_step + last_eval
    run.simulated_time = time_at(last_eval, lr) + (0.0 if final else float(lr[-1, 11]))
    run.wall_clock_s = time.time() - t_start
    run.resumed = resumed
    if checkpoint_path is not None:
        try:
            os.unlink(checkpoint_path)
        except FileNotFoundError:
            pass
    return run
