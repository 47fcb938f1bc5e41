"""Resident WCSPH stepper (libtsph ``tsph_sim_*``): the state lives in HBM
between steps; Python only queues work and reads meters.

``WcsphSimulation`` runs the loop of droplet_impact.run (droplet_impact.py:
1674-1763) / dam_break.run_case (dam_break.py:374-465): Verlet-skin pair
cache, rhs, symplectic Euler, the apply_boundary clamp, and per-step peak
meters, without a host round trip per step.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib


SimParams = _lib.SimParams
_vp, _i64 = C.c_void_p, C.c_int64

# "time" is the simulated time at the step's start (the sum of the step sizes
# before it), "last_dt" the previous step's size (op dt before the first),
# "step_dt" this step's size (0 for an adaptive run's final sample), "dt" the
# step-size variable the adaptive rule carries.
METER_KEYS = ["step", "mass", "momentum_x", "momentum_y", "momentum_z", "kinetic_energy",
              "rho_sum", "potential_energy", "rho_max", "pressure_max", "max_speed",
              "max_acceleration", "y_max", "x_max", "displacement", "peak_rho", "peak_speed",
              "rho_min", "pressure_min", "y_min", "time", "last_dt", "step_dt", "dt"]


def boundary_clamp(pos: torch.Tensor, vel: torch.Tensor, lo, hi, tol: float) -> int:
    """In-place clamp of FP64 (n, D) pos/vel; returns the crossing count."""
    dev = _lib.require_gpu(pos, vel)
    if pos.dtype != torch.float64 or not pos.is_contiguous():
        work_p = pos.detach().to(torch.float64).contiguous()
        work_v = vel.detach().to(torch.float64).contiguous()
    else:
        work_p, work_v = pos, vel
    dim = pos.shape[1]
    lo_a = (C.c_double * 3)(*(list(lo) + [0.0] * (3 - dim)))
    hi_a = (C.c_double * 3)(*(list(hi) + [0.0] * (3 - dim)))
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    _lib.check(_lib.lib().tsph_apply_boundary(_lib.ptr(work_p), _lib.ptr(work_v), pos.shape[0], dim,
                                              lo_a, hi_a, float(tol), _lib.ptr(cnt),
                                              _lib.stream_handle(dev)))
    if work_p is not pos:
        pos.copy_(work_p.to(pos.dtype))
        vel.copy_(work_v.to(vel.dtype))
    return int(cnt.item())


class WcsphSimulation:
    """A WCSPH simulation resident on one GPU.

    pos (n, D), vel (n, D), mass (n) are the fluid; wall (nw, D) static wall
    particles of mass ``wall_mass``.  ``precision`` is "fp64" or "fp32"
    (FP32 compensated).  ``lo``/``hi``/``tol`` give the apply_boundary box.
    ``adaptive`` (optional) = {"target_time", "courant_num", "dt_limit",
    "dt_safety", "dt_growth"[, "free_last", "tol", "initial_time",
    "initial_dt"]} runs the adaptive step of next_dt
    (droplet_impact.py:495-552) on the device up to target_time; ``dt`` is
    then the configured (largest) step.
    """

    def __init__(self, pos, vel, mass, wall, *, h, rho0, sound_speed, gamma, gravity, viscosity,
                 dt, wall_mass, skin, lo, hi, tol, wall_friction=0.0, density_diffusion=0.0,
                 diffusion_form="molteni", eigenvalue_floor=0.1, precision="fp32", clamp=True,
                 device="cuda", obstacle=None, adaptive=None, rhs_options=None):
        dev = torch.device(device)
        pos = torch.as_tensor(pos, dtype=torch.float64, device=dev)
        vel = torch.as_tensor(vel, dtype=torch.float64, device=dev)
        dim = pos.shape[1]
        wall = torch.zeros((0, dim), dtype=torch.float64, device=dev) if wall is None else \
            torch.as_tensor(wall, dtype=torch.float64, device=dev)
        self.n, self.nw, self.dim = pos.shape[0], wall.shape[0], dim
        self.code = _lib.F64 if precision == "fp64" else _lib.F32
        self.sdt = torch.float64 if self.code == _lib.F64 else torch.float32
        self.device = dev
        mass64 = torch.cat((torch.as_tensor(mass, dtype=torch.float64, device=dev),
                            torch.full((self.nw,), float(wall_mass), dtype=torch.float64,
                                       device=dev))).contiguous()
        mass_all = mass64.to(self.sdt).contiguous()
        pos_all = torch.cat((pos, wall)).contiguous()
        p = SimParams()
        op = p.op
        op.dim, op.precision, op.n_fluid, op.n_wall = dim, self.code, self.n, self.nw
        op.h, op.rho0, op.sound_speed, op.gamma = h, rho0, sound_speed, gamma
        op.gravity, op.viscosity, op.wall_friction, op.dt = gravity, viscosity, wall_friction, dt
        if density_diffusion > 0.0:
            op.diffusion_scale = 2.0 * density_diffusion * sound_speed * h
            op.diffusion_form = 1 if diffusion_form == "antuono" else 0
        op.eigenvalue_floor = eigenvalue_floor
        if rhs_options:
            # 2-D rhs switches (droplet_impact.py:954-964, 1031-1112): keys
            # density_renormalization + lattice_kernel_sum,
            # boundary_consistent_density, surface_pressure_closure +
            # surface_pressure_threshold / _donors / _layers
            if rhs_options.get("density_renormalization"):
                op.density_renormalization = 1
                op.lattice_kernel_sum = float(rhs_options["lattice_kernel_sum"])
            op.boundary_consistent_density = 1 if rhs_options.get("boundary_consistent_density") else 0
            if rhs_options.get("surface_pressure_closure"):
                op.surface_pressure_closure = 1
                op.surface_pressure_threshold = float(rhs_options["surface_pressure_threshold"])
                op.surface_pressure_donors_all = 1 if rhs_options["surface_pressure_donors"] == "all" else 0
                op.surface_pressure_layers = int(rhs_options["surface_pressure_layers"])
        p.skin = skin
        for d in range(3):
            p.lo[d] = lo[d] if d < dim else 0.0
            p.hi[d] = hi[d] if d < dim else 0.0
        p.tol = tol
        p.clamp = 1 if clamp else 0
        if obstacle is not None:
            # obstacle = {"rows": (begin, end) wall rows, "centre": (3,), optional
            # "box": (lo (3,), hi (3,)) for clamp_out_of_obstacle and "eps"}
            begin, end = obstacle["rows"]
            p.obstacle_begin, p.obstacle_end = int(begin), int(end)
            for d in range(3):
                p.obstacle_centre[d] = float(obstacle["centre"][d]) if d < len(obstacle["centre"]) else 0.0
            if obstacle.get("box") is not None:
                lo_b, hi_b = obstacle["box"]
                p.obstacle_clamp = 1
                for d in range(3):
                    p.obstacle_lo[d], p.obstacle_hi[d] = float(lo_b[d]), float(hi_b[d])
                p.obstacle_eps = float(obstacle.get("eps", 0.0))
        if adaptive is not None:
            p.adaptive = 1
            p.target_time = float(adaptive["target_time"])
            p.courant_num = float(adaptive["courant_num"])
            p.dt_limit = float(adaptive["dt_limit"])
            p.dt_safety = float(adaptive["dt_safety"])
            p.dt_growth = float(adaptive["dt_growth"])
            p.adaptive_free_last = 1 if adaptive.get("free_last") else 0
            p.adaptive_tol = float(adaptive.get("tol", 0.0))
            # a resumed run: the time and step-size variable it was saved with
            p.initial_time = float(adaptive.get("initial_time", 0.0))
            p.initial_dt = float(adaptive.get("initial_dt", 0.0))
        self.params = p
        handle = _vp()
        stream = _lib.stream_handle(dev)
        _lib.check(_lib.lib().tsph_sim_create(C.byref(p), _lib.ptr(pos_all), _lib.ptr(vel),
                                              _lib.ptr(mass_all), stream, C.byref(handle)))
        self.handle = handle
        if self.code == _lib.F32:
            # FP32 steps keep each mass as two floats and sum densities with
            # the FP64 value (the pressure contrast at depth H is ~h / H of p)
            _lib.check(_lib.lib().tsph_sim_set_masses_f64(handle, _lib.ptr(mass64), stream))
        self._keep = (pos_all, vel, mass_all, mass64)   # alive until the queued copies ran
        torch.cuda.current_stream(dev).synchronize()
        self._keep = None

    def step(self, nsteps: int, meter_every: int = 0) -> None:
        _lib.check(_lib.lib().tsph_sim_step(self.handle, int(nsteps), int(meter_every),
                                            _lib.stream_handle(self.device)))

    def read(self) -> dict:
        n, d = self.n, self.dim
        out = {
            "position": torch.empty((n, d), dtype=torch.float64, device=self.device),
            "velocity": torch.empty((n, d), dtype=torch.float64, device=self.device),
            "density": torch.empty(n, dtype=self.sdt, device=self.device),
            "pressure": torch.empty(n, dtype=self.sdt, device=self.device),
            "acceleration": torch.empty((n, d), dtype=self.sdt, device=self.device),
        }
        _lib.check(_lib.lib().tsph_sim_read(self.handle, _lib.ptr(out["position"]),
                                            _lib.ptr(out["velocity"]), _lib.ptr(out["density"]),
                                            _lib.ptr(out["pressure"]), _lib.ptr(out["acceleration"]),
                                            _lib.stream_handle(self.device)))
        return out

    def stats(self) -> dict:
        c = (_i64 * 6)()
        pk = (C.c_double * 2)()
        _lib.check(_lib.lib().tsph_sim_stats(self.handle, c, pk))
        return {"steps": c[0], "rebuilds": c[1], "escaped": c[2], "cells": c[3],
                "meter_rows": c[4], "finished": bool(c[5]), "peak_rho": pk[0], "peak_speed": pk[1]}

    def layout(self) -> dict:
        """Layout of the last build: "layout" (0 lists, 1 cell tiles, 2 lists
        with LDS halos), and for layout 2 the tiles of 256 rows on the halo
        path, all tiles and the largest halo in rows."""
        c = (_i64 * 4)()
        _lib.check(_lib.lib().tsph_sim_layout(self.handle, c))
        return {"layout": c[0], "halo_tiles": c[1], "tiles": c[2], "halo_max": c[3]}

    def slot_width(self, width: int | None = None) -> int:
        """The neighbour lists' slot width of the last build; a width sets the
        one the next build starts from (it widens by itself when a list does
        not fit)."""
        cur = _i64(0)
        _lib.check(_lib.lib().tsph_sim_slot_width(self.handle, -1 if width is None else int(width),
                                                  C.byref(cur)))
        return cur.value

    def time_state(self) -> dict:
        """The coming step's time state: simulated time at its start, the
        step-size variable, its own size, the previous step's size."""
        out = (C.c_double * 4)()
        _lib.check(_lib.lib().tsph_sim_time(self.handle, out, _lib.stream_handle(self.device)))
        return {"elapsed": out[0], "dt": out[1], "step_dt": out[2], "last_dt": out[3]}

    def meters(self) -> np.ndarray:
        nrows = _i64(0)
        _lib.check(_lib.lib().tsph_sim_meters(self.handle, None, 0, C.byref(nrows)))
        rows = np.zeros((nrows.value, len(METER_KEYS)), np.float64)
        if nrows.value:
            _lib.check(_lib.lib().tsph_sim_meters(
                self.handle, rows.ctypes.data_as(C.POINTER(C.c_double)), nrows.value,
                C.byref(nrows)))
        return rows

    def loads(self) -> np.ndarray:
        """Obstacle loads per completed step: (rows, 12) = step, F (3), M (3),
        cumulative clamp impulse (3), cumulative clamped particles, step size."""
        nrows = _i64(0)
        _lib.check(_lib.lib().tsph_sim_loads(self.handle, None, 0, C.byref(nrows)))
        rows = np.zeros((nrows.value, 12), np.float64)
        if nrows.value:
            _lib.check(_lib.lib().tsph_sim_loads(self.handle, rows.ctypes.data_as(C.POINTER(C.c_double)),
                                                 nrows.value, C.byref(nrows)))
        return rows

    def close(self) -> None:
        if getattr(self, "handle", None):
            _lib.lib().tsph_sim_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass



def _set_timing(self, enable, every: int = 1) -> None:
    """HIP-event sweep timing: on (of every ``every``-th step) or off; resets
    the sums."""
    _lib.check(_lib.lib().tsph_sim_timing(self.handle, max(int(every), 1) if enable else 0, None, None))


def _sweep_times(self):
    ms = (C.c_double * 3)()
    steps = _i64(0)
    _lib.check(_lib.lib().tsph_sim_timing(self.handle, -1, ms, C.byref(steps)))
    return [ms[0], ms[1], ms[2]], steps.value


def _configure(self, tiles: bool = True, tile=(8, 4), halo: bool = False) -> None:
    """Choose the sweep layout: LDS cell tiles (2-D only), or the per-row
    slot lists (tiles False), evaluated from the global arrays or, with halo
    True, from each workgroup's neighbourhood rows staged in LDS (2-D and
    3-D).  Every layout gives the same bits for the same lists."""
    if halo and tiles:
        raise ValueError("the halo layout runs on the slot lists: pass tiles=False")
    mode = 2 if halo else (1 if tiles else 0)
    _lib.check(_lib.lib().tsph_sim_configure(self.handle, mode, int(tile[0]), int(tile[1])))


def _count_pairs(self, radius: float) -> int:
    out = _i64(0)
    _lib.check(_lib.lib().tsph_sim_count_pairs(self.handle, float(radius), C.byref(out),
                                               _lib.stream_handle(self.device)))
    return out.value


def _gauges(self, spec) -> np.ndarray:
    """tsph_sim_gauges: the gauge row (GAUGE_ROW doubles) of the state the
    last step began from."""
    out = np.zeros(_lib.GAUGE_ROW, np.float64)
    _lib.check(_lib.lib().tsph_sim_gauges(self.handle, C.byref(spec),
                                          out.ctypes.data_as(C.POINTER(C.c_double)),
                                          _lib.stream_handle(self.device)))
    return out


def _shift(self, dt, v_ref=None, v_floor=0.0, strength=1.0, limit=0.25, free_surface=0.95,
           percentile=95.0, c_reference=None) -> dict:
    """tsph_sim_shift: one Lind shifting pass on the fluid after the last
    step (v_ref None: max(speed at the start of that step, v_floor); dt 0:
    that step's size; c_reference None: the percentile of the concentration).
    An explicit 0 is a value, as in the reference: v_ref = 0 moves nothing,
    c_reference = 0 marks no free surface."""
    if v_ref is not None and float(v_ref) < 0.0:
        raise ValueError("v_ref must be non-negative")
    if c_reference is not None and float(c_reference) < 0.0:
        raise ValueError("c_reference must be non-negative")
    p = _lib.ShiftParams(float(dt), -1.0 if v_ref is None else float(v_ref), float(v_floor),
                         float(strength), float(limit), float(free_surface), float(percentile),
                         -1.0 if c_reference is None else float(c_reference))
    info = (C.c_double * 4)()
    _lib.check(_lib.lib().tsph_sim_shift(self.handle, C.byref(p), info,
                                         _lib.stream_handle(self.device)))
    return {"c_reference": info[0], "max_shift": info[1], "clipped": info[2], "v_ref": info[3]}


def _tvc(self, alpha=0.05, limiter_slope=0.0, max_shift=None, wall_volume=0.0, scope="all",
         bulk_c=0.95, bulk_g=0.02, percentile=95.0) -> dict:
    """tsph_sim_tvc: one transport-velocity correction of the fluid after the
    last step (scope "all", "bulk" or "bulk_g"; max_shift None: no cap, 0:
    every shift clipped to zero, as in the reference)."""
    if max_shift is not None and float(max_shift) < 0.0:
        raise ValueError("max_shift must be non-negative")
    sc = {"bulk": 1, "bulk_g": 2}.get(scope, 0)
    p = _lib.TvcParams(float(alpha), float(limiter_slope),
                       -1.0 if max_shift is None else float(max_shift), float(wall_volume), sc,
                       float(bulk_c), float(bulk_g), float(percentile))
    info = (C.c_double * 3)()
    _lib.check(_lib.lib().tsph_sim_tvc(self.handle, C.byref(p), info, _lib.stream_handle(self.device)))
    return {"max_shift": info[0], "clipped": info[1], "crossings": int(info[2])}


WcsphSimulation.tvc = _tvc
WcsphSimulation.shift = _shift
WcsphSimulation.gauges = _gauges
WcsphSimulation.set_timing = _set_timing
WcsphSimulation.sweep_times = _sweep_times
WcsphSimulation.count_pairs = _count_pairs
WcsphSimulation.configure = _configure
