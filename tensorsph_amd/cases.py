"""Deterministic initial states of the benchmark cases (host-side set-up).

One-time lattice builders in float64, so the particle coordinates are the
reference's to the last bit.  The reference builds them with ``torch.arange``
and ``meshgrid``; ``_arange`` below calls ``torch.arange`` on the host, whose
values are not ``start + i * step`` rounded once (its vectorised fill rounds
some elements one ulp away), so the lattices match the reference's CPU runs
exactly.

* ``cell_centres``, ``make_particles``, ``wall_lattice``,
  ``lattice_kernel_sum``, ``lattice_mass_unit``, ``hydrostatic_density``:
  droplet_impact.py:559-754
* ``make_column``: dam_break.py:143-169
* ``make_column_3d``, ``wall_lattice_3d``, ``mass_3d``:
  dam_break_3d.py:172-283; 3-D lattice constant threed/kernel3.py:184-215
* ``obstacle_lattice_3d``: kleefsman_obstacle_3d.py:187-205
* ``cubic_lattice``: threed/solver.py:84-90
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np
import torch

from .config import Config, DamBreakConfig, DamBreak3DConfig, Kleefsman3DConfig


def _arange(start: float, end: float, step: float) -> np.ndarray:
    return torch.arange(float(start), float(end), float(step), dtype=torch.float64).numpy()


# ---------------------------------------------------------------- kernels (host)

def cubic_w2(r: np.ndarray, h: float) -> np.ndarray:
    """2-D cubic spline W (droplet_impact.py:276-288), host-side, for set-up sums."""
    q = r / h
    sigma = 10.0 / (7.0 * math.pi * h * h)
    first = 1.0 - 1.5 * q ** 2 + 0.75 * q ** 3
    second = 0.25 * np.clip(2.0 - q, 0.0, None) ** 3
    shape = np.where(q < 1.0, first, np.where(q < 2.0, second, 0.0))
    return sigma * np.clip(shape, 0.0, None)


def cubic_w3(r: np.ndarray, h: float) -> np.ndarray:
    """3-D cubic spline W (threed/kernel3.py:95-102), host-side."""
    q = r / h
    sigma = 1.0 / (math.pi * h * h * h)
    first = 1.0 - 1.5 * q ** 2 + 0.75 * q ** 3
    second = 0.25 * np.clip(2.0 - q, 0.0, None) ** 3
    shape = np.where(q < 1.0, first, np.where(q < 2.0, second, 0.0))
    return sigma * np.clip(shape, 0.0, None)


def lattice_kernel_sum(h: float, dx: float) -> float:
    """Discrete 2-D kernel sum over a full square lattice (droplet_impact.py:577-598)."""
    if h <= 0.0 or dx <= 0.0:
        raise ValueError("h and dx must be positive")
    reach = int(math.ceil(2.0 * h / dx)) + 1
    offsets = np.arange(-reach, reach + 1, dtype=np.float64) * dx
    gx, gy = np.meshgrid(offsets, offsets, indexing="xy")
    r = np.sqrt(gx ** 2 + gy ** 2).reshape(-1)
    return float(cubic_w2(r, h).sum())


def lattice_mass_unit(cfg: Config) -> float:
    """droplet_impact.lattice_mass_unit (:618-622): calibrated ``rho0 / S`` or ``rho0 dx^2``."""
    if cfg.density_calibration:
        return float(cfg.rho0 / lattice_kernel_sum(cfg.h_factor * cfg.dx, cfg.dx))
    return cfg.rho0 * cfg.dx * cfg.dx


def lattice_kernel_sum_3d(dx: float, h: float, half_width: int | None = None) -> float:
    """``dx^3 sum W`` over a simple cubic lattice (threed/kernel3.py:184-197)."""
    if half_width is None:
        half_width = int(math.ceil(2.0 * h / dx)) + 1
    n = np.arange(-half_width, half_width + 1, dtype=np.float64) * dx
    g = np.stack(np.meshgrid(n, n, n, indexing="ij"), -1).reshape(-1, 3)
    r = np.sqrt((g ** 2).sum(1))
    return float(cubic_w3(r, h).sum() * dx ** 3)


def calibrated_particle_mass_3d(dx: float, h: float, rho0: float) -> float:
    """threed/kernel3.py:206-215."""
    return rho0 * dx ** 3 / lattice_kernel_sum_3d(dx, h)


@lru_cache(maxsize=64)
def mass_3d(dx: float, h_factor: float, rho0: float, calibrated: bool) -> float:
    """dam_break_3d._mass_3d (dam_break_3d.py:172-178)."""
    if calibrated:
        return calibrated_particle_mass_3d(dx, h_factor * dx, rho0)
    return rho0 * dx ** 3


# ---------------------------------------------------------------- 2-D lattices

def domain_top(cfg: Config) -> float:
    return cfg.drop_center_y + cfg.drop_radius + 2.0 * cfg.dx


def hydrostatic_density(y: np.ndarray, cfg: Config, surface: float | None = None) -> np.ndarray:
    """droplet_impact.hydrostatic_density (:565-575)."""
    bulk = cfg.rho0 * cfg.sound_speed ** 2 / cfg.gamma
    top = cfg.pool_depth if surface is None else surface
    depth = np.clip(top - y, 0.0, None)
    return cfg.rho0 * (1.0 + cfg.rho0 * cfg.gravity * depth / bulk) ** (1.0 / cfg.gamma)


def cell_centres(extent: float, dx: float) -> np.ndarray:
    ratio = float(extent) / float(dx)
    count = int(math.floor(ratio + 1.0e-9))
    if count < 1:
        raise ValueError(f"dx={dx:g} is too coarse for an extent of {extent:g} "
                         f"(only {ratio:.3f} cells fit)")
    return (np.arange(count, dtype=np.float64) + 0.5) * dx


def make_particles(cfg: Config):
    """Pool + drop lattice and masses (droplet_impact.py:666-708)."""
    dx = cfg.dx
    if cfg.drop_center_y - cfg.drop_radius <= cfg.pool_depth:
        raise ValueError("drop must start above the pool surface")
    xs = cell_centres(cfg.pool_width, dx)
    ys = cell_centres(cfg.pool_depth, dx)
    gx, gy = np.meshgrid(xs, ys, indexing="xy")
    pool = np.stack((gx.reshape(-1), gy.reshape(-1)), -1)
    drop_y = _arange(cfg.drop_center_y - cfg.drop_radius, cfg.drop_center_y + cfg.drop_radius, dx)
    if drop_y.size == 0:
        raise ValueError("drop_radius is too small for the particle spacing")
    dxg, dyg = np.meshgrid(xs, drop_y, indexing="xy")
    cand = np.stack((dxg.reshape(-1), dyg.reshape(-1)), -1)
    centre = np.array([cfg.pool_width * 0.5, cfg.drop_center_y])
    drop = cand[np.sqrt(((cand - centre) ** 2).sum(1)) <= cfg.drop_radius]
    if drop.shape[0] == 0:
        raise ValueError("drop discretization produced no particles")
    pos = np.concatenate((pool, drop))
    unit = lattice_mass_unit(cfg)
    mass = np.full(pos.shape[0], unit)
    if cfg.hydrostatic_init:
        mass[: pool.shape[0]] = hydrostatic_density(pool[:, 1], cfg) * (unit / cfg.rho0)
    return pos, mass, pool.shape[0]


def wall_lattice(cfg: Config, top: float | None = None) -> np.ndarray:
    """Three-layer static wall on bottom and sides (droplet_impact.py:711-744)."""
    if cfg.wall_layers < 1:
        raise ValueError("wall_layers must be positive")
    dx, layers = cfg.dx, cfg.wall_layers
    top = domain_top(cfg) if top is None else float(top)
    span_x = _arange(-(layers - 0.5) * dx, cfg.pool_width + (layers - 0.5) * dx, dx)
    columns = _arange(dx * 0.5, top, dx)
    parts = []
    for layer in range(1, layers + 1):
        parts.append(np.stack((span_x, np.full_like(span_x, dx * 0.5 - layer * dx)), -1))
    for layer in range(1, layers + 1):
        for x in (dx * 0.5 - layer * dx, cfg.pool_width - dx * 0.5 + layer * dx):
            parts.append(np.stack((np.full_like(columns, x), columns), -1))
    return np.concatenate(parts)


def make_column(cfg: DamBreakConfig):
    """Dam-break water column (dam_break.py:143-169)."""
    xs = cell_centres(cfg.column_width, cfg.dx)
    ys = cell_centres(cfg.column_height, cfg.dx)
    gx, gy = np.meshgrid(xs, ys, indexing="xy")
    pos = np.stack((gx.reshape(-1), gy.reshape(-1)), -1)
    unit = lattice_mass_unit(cfg)
    if cfg.hydrostatic_init:
        mass = hydrostatic_density(pos[:, 1], cfg, surface=cfg.column_height) * (unit / cfg.rho0)
    else:
        mass = np.full(pos.shape[0], unit)
    return pos, mass


def impact_state(**overrides):
    """Config + pool/drop + walls of the droplet-impact case, as numpy arrays."""
    cfg = Config(device="cpu", **overrides)
    pos, mass, _ = make_particles(cfg)
    wall = wall_lattice(cfg, top=cfg.wall_top)
    return cfg, pos, mass, wall


# ---------------------------------------------------------------- 3-D lattices

def make_column_3d(cfg: DamBreak3DConfig):
    """3-D column (dam_break_3d.py:185-220), y up, ordered y, x, z (z fastest)."""
    dx = cfg.dx
    xs = cell_centres(cfg.column_width, dx)
    ys = cell_centres(cfg.column_height, dx)
    zs = cell_centres(cfg.column_depth, dx)
    gx, gy, gz = np.meshgrid(xs, ys, zs, indexing="xy")
    pos = np.stack((gx.reshape(-1), gy.reshape(-1), gz.reshape(-1)), -1)
    unit = mass_3d(dx, cfg.h_factor, cfg.rho0, cfg.density_calibration)
    if cfg.hydrostatic_init:
        mass = hydrostatic_density(pos[:, 1], cfg, surface=cfg.column_height) * (unit / cfg.rho0)
    else:
        mass = np.full(pos.shape[0], unit)
    return pos, mass


def wall_lattice_3d(cfg: DamBreak3DConfig, top: float | None = None) -> np.ndarray:
    """Bottom and four side walls (dam_break_3d.py:223-283)."""
    if cfg.wall_layers < 1:
        raise ValueError("wall_layers must be positive")
    dx, layers = cfg.dx, cfg.wall_layers
    top = cfg.tank_top() if top is None else float(top)
    tw, td = cfg.pool_width, cfg.tank_depth
    span_x = _arange(-(layers - 0.5) * dx, tw + (layers - 0.5) * dx, dx)
    span_z = _arange(-(layers - 0.5) * dx, td + (layers - 0.5) * dx, dx)
    cols_y = _arange(dx * 0.5, top, dx)
    span_x_wall = _arange(dx * 0.5, tw, dx)
    parts = []

    def grid(a, b, c):
        ga, gb, gc = np.meshgrid(a, b, c, indexing="xy")
        return np.stack((ga.reshape(-1), gb.reshape(-1), gc.reshape(-1)), -1)

    for layer in range(1, layers + 1):
        parts.append(grid(span_x, np.array([dx * 0.5 - layer * dx]), span_z))
    for layer in range(1, layers + 1):
        parts.append(grid(np.array([dx * 0.5 - layer * dx]), cols_y, span_z))
    for layer in range(1, layers + 1):
        parts.append(grid(np.array([tw - dx * 0.5 + layer * dx]), cols_y, span_z))
    for layer in range(1, layers + 1):
        parts.append(grid(span_x_wall, cols_y, np.array([dx * 0.5 - layer * dx])))
    for layer in range(1, layers + 1):
        parts.append(grid(span_x_wall, cols_y, np.array([td - dx * 0.5 + layer * dx])))
    return np.concatenate(parts)


def obstacle_lattice_3d(cfg: Kleefsman3DConfig) -> np.ndarray:
    """Box on the floor (kleefsman_obstacle_3d.py:187-205).

    ``cfg.obstacle_fill == "filled"`` is the reference's filled block; "shell"
    keeps the particles within ``wall_layers`` spacings of a face."""
    dx = cfg.dx

    def span(lo, hi):
        n = max(1, int(round((hi - lo) / dx)))
        return lo + (np.arange(n, dtype=np.float64) + 0.5) * dx

    xs = span(cfg.obstacle_x, cfg.obstacle_x1())
    ys = span(0.0, cfg.obstacle_height)
    zs = span(cfg.obstacle_z, cfg.obstacle_z1())
    gx, gy, gz = np.meshgrid(xs, ys, zs, indexing="xy")
    pts = np.stack((gx.reshape(-1), gy.reshape(-1), gz.reshape(-1)), -1)
    if getattr(cfg, "obstacle_fill", "filled") == "shell":
        ix, iy, iz = np.meshgrid(np.arange(len(xs)), np.arange(len(ys)), np.arange(len(zs)),
                                 indexing="xy")
        L = cfg.wall_layers
        depth = np.minimum.reduce([ix, len(xs) - 1 - ix, iy, len(ys) - 1 - iy,
                                   iz, len(zs) - 1 - iz]).reshape(-1)
        # the floor face rests on the tank bottom wall: no shell needed there
        depth_nofloor = np.minimum.reduce([ix, len(xs) - 1 - ix, len(ys) - 1 - iy,
                                           iz, len(zs) - 1 - iz]).reshape(-1)
        pts = pts[depth_nofloor < L]
        del depth
    return pts


def cubic_lattice(shape, dx: float, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """threed/solver.py:84-90: points at ``i * dx + origin``, row-major (z fastest)."""
    axes = [np.arange(n, dtype=np.float64) * dx + origin[a] for a, n in enumerate(shape)]
    g = np.meshgrid(*axes, indexing="ij")
    return np.stack(g, -1).reshape(-1, len(shape))
