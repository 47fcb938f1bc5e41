"""3-D dam break against a box, behind the reference's ``kleefsman_obstacle_3d``
API (kleefsman_obstacle_3d.py).

The driver, run record, gauges and checkpoints live in ``kleefsman``; this
module gives them the reference's module name and adds what its call sites use
directly: ``validate_case`` (:129-179), ``obstacle_lattice_3d(cfg, device)``
(:187-205), ``clamp_out_of_obstacle`` (:208-266, on the GPU through
``tsph_obstacle_clamp``) and ``score`` (:733-785).
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from . import _lib
from . import cases
from .config import (COLUMN_HEIGHT_M, COLUMN_LENGTH_M, FROZEN_APERTURE_M,  # noqa: F401
                     FROZEN_ARRIVAL_THRESHOLD_M, FROZEN_STATION_OFFSET_M, H2_X_M,
                     OBSTACLE_HEIGHT_M, OBSTACLE_LENGTH_M, OBSTACLE_WIDTH_M, OBSTACLE_X_M,
                     TANK_HEIGHT_M, TANK_LENGTH_M, TANK_WIDTH_M, Kleefsman3DConfig)
from .kleefsman import (CHECKPOINT_FINGERPRINT_FIELDS, CHECKPOINT_VERSION,  # noqa: F401
                        H2_ARRIVAL_S, H2_PEAK_M, H2_PEAK_TIME_S, Kleefsman3DRun, arrival_time,
                        bypass_count, checkpoint_fingerprint, load_checkpoint, peak_height,
                        run_case, save_checkpoint, slab_height)

REFERENCE_PEAK_HEIGHT_M = H2_PEAK_M
REFERENCE_PEAK_TIME_S = H2_PEAK_TIME_S
REFERENCE_ARRIVAL_TIME_S = H2_ARRIVAL_S
TOLERANCE_REL = 0.10

__all__ = [
    "Kleefsman3DConfig", "Kleefsman3DRun", "arrival_time", "bypass_count",
    "clamp_out_of_obstacle", "obstacle_lattice_3d", "peak_height", "run_case",
    "score", "slab_height", "validate_case", "save_checkpoint", "load_checkpoint",
    "checkpoint_fingerprint",
    "REFERENCE_PEAK_HEIGHT_M", "REFERENCE_PEAK_TIME_S",
    "REFERENCE_ARRIVAL_TIME_S", "TOLERANCE_REL",
]


def validate_case(cfg: Kleefsman3DConfig) -> None:
    """Reject geometry that is not the published apparatus
    (kleefsman_obstacle_3d.py:129-179), with the reference's messages."""
    if not math.isfinite(cfg.viscosity) or cfg.viscosity < 0:
        raise ValueError("kinematic viscosity must be finite and non-negative")
    tol = 1.0e-9
    checks = (
        (cfg.pool_width, TANK_LENGTH_M, "tank length must be 3.22 m (Kleefsman et al. 2005)"),
        (cfg.tank_height, TANK_HEIGHT_M, "tank height must be 1.00 m"),
        (cfg.tank_depth, TANK_WIDTH_M, "tank width must be 1.00 m"),
        (cfg.column_width, COLUMN_LENGTH_M, "column length must be 1.228 m"),
        (cfg.column_height, COLUMN_HEIGHT_M, "column height must be 0.550 m"),
        (cfg.column_depth, cfg.tank_depth, "the water column must span the full tank width"),
        (cfg.obstacle_x, OBSTACLE_X_M, "obstacle upstream face must be at x = 2.3955 m"),
        (cfg.obstacle_height, OBSTACLE_HEIGHT_M, "obstacle height must be 0.161 m"),
        (cfg.obstacle_width, OBSTACLE_WIDTH_M, "obstacle width must be 0.403 m (3-D case only)"),
    )
    for got, want, msg in checks:
        if abs(got - want) > tol:
            raise ValueError(msg)
    if cfg.obstacle_z < -tol or cfg.obstacle_z1() > cfg.tank_depth + tol:
        raise ValueError("obstacle must lie inside the tank width")
    if cfg.obstacle_x1() > cfg.pool_width:
        raise ValueError("obstacle must lie inside the tank length")
    if cfg.obstacle_height > cfg.tank_height:
        raise ValueError("obstacle must be shorter than the tank")
    if cfg.dx <= 0.0:
        raise ValueError("dx must be positive")
    for name, extent in (("length", cfg.obstacle_length), ("height", cfg.obstacle_height),
                         ("width", cfg.obstacle_width)):
        if extent < 3.0 * cfg.dx:
            raise ValueError(f"obstacle {name} {extent:g} m is under-resolved at dx={cfg.dx:g}")
    if cfg.gauge_aperture <= 0.0 or cfg.gauge_aperture >= cfg.obstacle_x:
        raise ValueError("gauge aperture must be positive and upstream of the box")
    if cfg.station_x() <= cfg.column_width:
        raise ValueError("gauge station falls inside the reservoir")
    courant = cfg.dt * cfg.sound_speed / (cfg.h_factor * cfg.dx)
    if courant > cfg.max_courant * (1.0 + 1.0e-9):
        raise ValueError(
            f"dt={cfg.dt:g} gives acoustic Courant {courant:.3f} > "
            f"max_courant {cfg.max_courant:g}; use dt <= {cfg.dt_from_courant():.3e}")


def obstacle_lattice_3d(cfg: Kleefsman3DConfig, device) -> torch.Tensor:
    """Solid particles filling the box on the fluid lattice, float64
    (kleefsman_obstacle_3d.py:187-205)."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError(f"tensorsph_amd runs on the GPU: pass a 'cuda' (ROCm) device (got {dev})")
    return torch.as_tensor(cases.obstacle_lattice_3d(cfg), dtype=torch.float64, device=dev)


def clamp_out_of_obstacle(pos: torch.Tensor, vel: torch.Tensor, cfg: Kleefsman3DConfig, *,
                          mass=None, body_impulse=None) -> int:
    """Move fluid particles strictly inside the box out through the nearest
    face (to 0.01 dx beyond it) and zero that velocity component, in place;
    returns how many were moved (kleefsman_obstacle_3d.py:208-266).  With
    ``body_impulse`` (a (3,) tensor) the momentum removed, sum m (v_before -
    v_after), is added to it; that needs ``mass``."""
    if pos.numel() == 0:
        return 0
    if pos.ndim != 2 or pos.shape[1] != 3 or vel.shape != pos.shape:
        raise ValueError("pos and vel must be (N, 3)")
    if body_impulse is not None and mass is None:
        raise TypeError("body_impulse needs the particle masses (mass=...)")
    dev = _lib.require_gpu(pos, vel, mass, body_impulse)
    direct = pos.dtype == torch.float64 and pos.is_contiguous() and \
        vel.dtype == torch.float64 and vel.is_contiguous()
    work_p = pos if direct else pos.detach().to(torch.float64).contiguous()
    work_v = vel if direct else vel.detach().to(torch.float64).contiguous()
    lo = (C.c_double * 3)(cfg.obstacle_x, 0.0, cfg.obstacle_z)
    hi = (C.c_double * 3)(cfg.obstacle_x1(), cfg.obstacle_height, cfg.obstacle_z1())
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    m = imp = None
    if body_impulse is not None:
        m = mass.detach().to(torch.float64).contiguous()
        imp = torch.zeros(3, dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().tsph_obstacle_clamp(_lib.ptr(work_p), _lib.ptr(work_v), _lib.ptr(m),
                                              _lib.F64, pos.shape[0], lo, hi, 0.01 * cfg.dx,
                                              _lib.ptr(cnt), _lib.ptr(imp),
                                              _lib.stream_handle(dev)))
    count = int(cnt.item())
    if not direct and count:
        pos.copy_(work_p.to(pos.dtype))
        vel.copy_(work_v.to(vel.dtype))
    if imp is not None:
        body_impulse.add_(imp.to(body_impulse.dtype))
    return count


def score(run: Kleefsman3DRun, tolerance_rel: float = TOLERANCE_REL) -> dict:
    """Peak and arrival of the height in front of the box against the SPHERIC
    workbook values (kleefsman_obstacle_3d.py:733-785)."""
    peak_h, _ = run.peak()
    arrival = run.arrival()
    rows = []

    def add(metric, ref, got):
        rel = abs(got - ref) / ref if got is not None and ref else float("inf")
        rows.append({"metric": metric, "ref": ref, "got": got, "rel_error": rel,
                     "pass": bool(rel <= tolerance_rel)})

    add("peak_height", REFERENCE_PEAK_HEIGHT_M, peak_h)
    if arrival is None:
        rows.append({"metric": "arrival_time", "ref": REFERENCE_ARRIVAL_TIME_S, "got": None,
                     "rel_error": float("inf"), "pass": False})
    else:
        add("arrival_time", REFERENCE_ARRIVAL_TIME_S, arrival)
    censored = not run.times or run.times[-1] < REFERENCE_PEAK_TIME_S
    rows[0]["assessable"] = not censored
    if censored:
        rows[0]["pass"] = False
    has_local = bool(run.local_h2) and len(run.local_h2) == len(run.times)
    add("H2_local_peak", REFERENCE_PEAK_HEIGHT_M, max(run.local_h2) if has_local else None)
    add("H2_local_arrival", REFERENCE_ARRIVAL_TIME_S,
        arrival_time(run.times, run.local_h2, run.cfg.arrival_threshold) if has_local else None)
    rows[-2]["assessable"] = has_local and not censored
    if not rows[-2]["assessable"]:
        rows[-2]["pass"] = False
    passed = all(r["pass"] for r in rows) and not censored
    return {
        "case": "CASE-SPH-KLEEFSMAN-DAMBREAK-OBST",
        "assessment_scope": ("SPHERIC workbook H2, local and blocked-width volume gauges; "
                             "water level only"),
        "geometry": "3-D (0.403 m box in 1.00 m tank, 40 % blockage)",
        "reference": {
            "peak_height_m": REFERENCE_PEAK_HEIGHT_M,
            "peak_time_s": REFERENCE_PEAK_TIME_S,
            "arrival_time_s": REFERENCE_ARRIVAL_TIME_S,
            "source": "SPHERIC Test 2, test_case_2_exp_data.xls, Experimental_data / H2 (m)",
        },
        "tolerance_rel": tolerance_rel,
        "rows": rows,
        "window_covers_reference_peak": bool(run.times and run.times[-1] >= REFERENCE_PEAK_TIME_S),
        "censored": censored,
        "verdict": "PASS" if passed else "FAIL",
        "run": run.to_dict(),
    }
