"""ctypes front end of the CPU oracle (``tsph_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke``
and by bench.py's CPU-baseline leg, never by the product package.  Arrays are
numpy, row-major (AoS), float64 unless a ``_f32`` entry point is named.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libtsph_oracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_lp = C.POINTER(C.c_int64)


class MorrisParams(C.Structure):
    _fields_ = [("dim", C.c_int), ("kernel3d", C.c_int),
                ("h", C.c_double), ("rho0", C.c_double), ("c0", C.c_double),
                ("gamma", C.c_double), ("gravity", C.c_double), ("nu", C.c_double),
                ("wall_friction", C.c_double), ("dt", C.c_double),
                ("diff_scale", C.c_double), ("diff_form", C.c_int),
                ("eig_floor", C.c_double), ("reaction_style", C.c_int),
                ("wrap_width", C.c_double), ("shepard", C.c_int), ("lattice_sum", C.c_double),
                ("bcd", C.c_int), ("closure", C.c_int), ("closure_threshold", C.c_double),
                ("closure_literal", C.c_int), ("closure_layers", C.c_int),
                ("density_offset", C.c_void_p), ("nu_field", C.c_void_p)]


class DPParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("dx", "c0", "ref_speed", "dt", "rho0", "gravity",
                                          "nu", "diffusion", "alpha")]


def _declare(L):
    for suf, ptr in (("f64", _dp), ("f32", _fp)):
        f = getattr(L, f"orc_cell_linked_pairs_{suf}")
        f.restype = C.c_int64
        f.argtypes = [ptr, C.c_int64, C.c_int, C.c_double, C.POINTER(_lp), _lp, _lp, _lp]
        f = getattr(L, f"orc_brute_force_pairs_{suf}")
        f.restype = C.c_int64
        f.argtypes = [ptr, C.c_int64, C.c_int, C.c_double, C.POINTER(_lp)]
    L.orc_periodic_pairs_f64.restype = C.c_int64
    L.orc_periodic_pairs_f64.argtypes = [_dp, C.c_int64, C.c_int, _dp, C.c_double, C.POINTER(_lp)]
    L.orc_free.argtypes = [C.c_void_p]
    L.orc_kernel_eval.argtypes = [C.c_int, _dp, C.c_int64, C.c_double, _dp, _dp, _dp]
    L.orc_sym_eig.argtypes = [C.c_int, _dp, _dp, _dp]
    L.orc_tait2d.argtypes = [_dp, C.c_int64, C.c_double, C.c_double, C.c_double, _dp]
    L.orc_rhs_morris.argtypes = [C.POINTER(MorrisParams), C.c_int64, C.c_int64, _dp, _dp, _dp,
                                 _dp, _lp, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_threed_accelerations.argtypes = [C.c_int64, C.c_int64, _dp, _dp, _dp, C.c_double, _lp,
                                           C.c_int64, C.c_double, C.c_double, C.c_double,
                                           C.c_double, _dp, _dp, C.c_int, C.c_double,
                                           C.c_double, C.c_int, C.c_double, C.c_int,
                                           _dp, _dp, _dp]
    L.orc_apply_boundary.restype = C.c_int64
    L.orc_apply_boundary.argtypes = [C.c_int, C.c_int64, _dp, _dp, _dp, _dp, C.c_double]
    L.orc_delta_plus_step.restype = C.c_int
    L.orc_delta_plus_step.argtypes = [C.POINTER(DPParams), C.c_int64, C.c_int64, _dp, _dp, _dp,
                                      _dp, _dp, _dp, _dp, _lp, C.c_int64, _dp, _dp, _dp, _dp,
                                      _dp, _dp]


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _l(a):
    return a.ctypes.data_as(_lp) if a is not None else None


def _take_pairs(ptr, count):
    out = np.ctypeslib.as_array(ptr, shape=(max(count, 0) * 2,)).copy().reshape(-1, 2) \
        if count > 0 else np.zeros((0, 2), np.int64)
    lib().orc_free(ptr)
    return out


def cell_linked_pairs(pos, cutoff, with_cells=False):
    """neighbor_search.cell_linked_pairs (neighbor_search.py:44-120)."""
    pos = np.ascontiguousarray(pos)
    n, dim = pos.shape
    f64 = pos.dtype == np.float64
    if not f64:
        pos = pos.astype(np.float32)
    cells = np.zeros((n, dim), np.int64)
    keys = np.zeros(n, np.int64)
    order = np.zeros(n, np.int64)
    out = _lp()
    fn = lib().orc_cell_linked_pairs_f64 if f64 else lib().orc_cell_linked_pairs_f32
    data = pos.ctypes.data_as(_dp if f64 else _fp)
    cnt = fn(data, n, dim, float(cutoff), C.byref(out), _l(cells), _l(keys), _l(order))
    pairs = _take_pairs(out, cnt)
    if with_cells:
        return pairs, cells, keys, order
    return pairs


def brute_force_pairs(pos, cutoff):
    pos = np.ascontiguousarray(pos)
    n, dim = pos.shape
    f64 = pos.dtype == np.float64
    if not f64:
        pos = pos.astype(np.float32)
    out = _lp()
    fn = lib().orc_brute_force_pairs_f64 if f64 else lib().orc_brute_force_pairs_f32
    cnt = fn(pos.ctypes.data_as(_dp if f64 else _fp), n, dim, float(cutoff), C.byref(out))
    return _take_pairs(out, cnt)


def periodic_pairs(pos, box, cutoff):
    """threed.neighbors3.periodic_pairs (neighbors3.py:71-142)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n, dim = pos.shape
    b = np.broadcast_to(np.asarray(box, np.float64), (dim,)).copy()
    out = _lp()
    cnt = lib().orc_periodic_pairs_f64(_d(pos), n, dim, _d(b), float(cutoff), C.byref(out))
    return _take_pairs(out, cnt)


def kernel_eval(kind, r, h):
    """kind 0: 2-D cubic, 1: 3-D cubic (+Laplacian), 2: 2-D Wendland C2."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    w = np.zeros_like(r); s = np.zeros_like(r); lap = np.zeros_like(r)
    lib().orc_kernel_eval(kind, _d(r), r.size, float(h), _d(w), _d(s), _d(lap))
    return w, s, lap


def sym_eig(m):
    m = np.ascontiguousarray(m, dtype=np.float64)
    n = m.shape[0]
    w = np.zeros(n); v = np.zeros((n, n))
    lib().orc_sym_eig(n, _d(m), _d(w), _d(v))
    return w, v


def tait2d(rho, rho0, c0, gamma):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    p = np.zeros_like(rho)
    lib().orc_tait2d(_d(rho), rho.size, rho0, c0, gamma, _d(p))
    return p


def rhs_morris(pos, vel, mass, wall, pairs, *, dim, h, rho0, c0, gamma, gravity, nu,
               wall_friction=0.0, dt=0.0, diff_scale=0.0, diff_form=0, eig_floor=0.1,
               wall_mass=None, wall_vel=None, wall_accel=None, kernel3d=False,
               reaction_style=0, wrap_width=0.0, shepard=False, lattice_sum=0.0, bcd=False,
               closure=False, closure_threshold=0.99, closure_literal=False, closure_layers=2,
               density_offset=None, nu_field=None):
    """droplet_impact.rhs (2-D) / dam_break_3d.rhs_3d (3-D) in FP64, with the
    2-D rhs's opt-in branches: wrap_width, density_offset, Shepard density
    renormalisation (lattice_sum = lattice_kernel_sum(h, dx)), the
    surface-pressure closure, boundary-consistent density and nu_field.

    Returns (rho, p, acc, wall_p, wall_rho, reaction)."""
    pos = np.asarray(pos, np.float64); vel = np.asarray(vel, np.float64)
    mass = np.asarray(mass, np.float64)
    n = pos.shape[0]
    wall = np.zeros((0, dim)) if wall is None else np.asarray(wall, np.float64)
    nw = wall.shape[0]
    pos_all = np.ascontiguousarray(np.concatenate([pos, wall]))
    wv = np.zeros((nw, dim)) if wall_vel is None else np.asarray(wall_vel, np.float64)
    vel_all = np.ascontiguousarray(np.concatenate([vel, wv]))
    wm = np.full(nw, 0.0 if wall_mass is None else float(wall_mass))
    mass_all = np.ascontiguousarray(np.concatenate([mass, wm]))
    pairs = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1, 2)
    off = None if density_offset is None else np.ascontiguousarray(density_offset, np.float64)
    nuf = None if nu_field is None else np.ascontiguousarray(nu_field, np.float64)
    P = MorrisParams(dim, 1 if kernel3d else 0, h, rho0, c0, gamma, gravity, nu,
                     wall_friction, dt, diff_scale, diff_form, eig_floor, reaction_style,
                     float(wrap_width), int(bool(shepard)), float(lattice_sum), int(bool(bcd)),
                     int(bool(closure)), float(closure_threshold), int(bool(closure_literal)),
                     int(closure_layers), None if off is None else off.ctypes.data,
                     None if nuf is None else nuf.ctypes.data)
    rho = np.zeros(n); p = np.zeros(n); acc = np.zeros((n, dim))
    wp = np.zeros(nw); wr = np.zeros(nw); reac = np.zeros((nw, dim))
    wa = None if wall_accel is None else np.ascontiguousarray(wall_accel, np.float64)
    lib().orc_rhs_morris(C.byref(P), n, nw, _d(pos_all), _d(vel_all), _d(mass_all), _d(wa),
                         _l(pairs), pairs.shape[0], _d(rho), _d(p), _d(acc), _d(wp), _d(wr),
                         _d(reac) if reaction_style else None)
    return rho, p, acc, wp, wr, reac


def threed_accelerations(pos, vel, mass, h, pairs, *, rho0, sound_speed, gamma=7.0,
                         viscosity=0.0, gravity=None, box=None, n_fluid=None,
                         renormalise=True, density_diffusion=0.0, dt=None,
                         density_diffusion_form="antuono",
                         density_diffusion_eigenvalue_floor=0.1, viscous_order=2):
    """threed.operators3.accelerations (operators3.py:212-256), FP64, with the
    viscous operator of order 2 or 4 (operators3.py:116-154)."""
    if viscous_order not in (2, 4):
        raise ValueError(f"viscous_order must be 2 or 4, got {viscous_order}")
    pos = np.ascontiguousarray(pos, np.float64); vel = np.ascontiguousarray(vel, np.float64)
    mass = np.ascontiguousarray(mass, np.float64)
    n = pos.shape[0]
    nf = n if n_fluid is None else int(n_fluid)
    pairs = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1, 2)
    g = None if gravity is None else np.ascontiguousarray(gravity, np.float64)
    b = None if box is None else np.broadcast_to(np.asarray(box, np.float64), (3,)).copy()
    acc = np.zeros((n, 3)); rho = np.zeros(n); p = np.zeros(n)
    form = 1 if density_diffusion_form == "antuono" else 0
    lib().orc_threed_accelerations(n, nf, _d(pos), _d(vel), _d(mass), h, _l(pairs),
                                   pairs.shape[0], rho0, sound_speed, gamma, viscosity,
                                   _d(g), _d(b), 1 if renormalise else 0,
                                   float(density_diffusion), float(dt or 0.0), form,
                                   density_diffusion_eigenvalue_floor, int(viscous_order),
                                   _d(acc), _d(rho), _d(p))
    return acc, rho, p


def ledger_row(pos, vel, mass, rho, *, rho0, sound_speed, gamma, box, wavenumber):
    """Conservation row (threed/conservation.py:106-121) plus the shear-wave
    projections (operators3.py:401-408): mass, momentum, angular momentum,
    kinetic and Tait internal energy, sum u_x sin(kz), sum u_x sin(ky) sin(kz),
    sum m |v|_1."""
    m = mass[:, None]
    ratio = rho / rho0
    e = (sound_speed ** 2 / gamma) * ((ratio ** (gamma - 1.0) - 1.0) / (gamma - 1.0)
                                      + 1.0 / ratio - 1.0)
    return {"mass": float(mass.sum()), "momentum": (m * vel).sum(0),
            "angular_momentum": (m * np.cross(pos, vel)).sum(0),
            "kinetic_energy": float(0.5 * (m * vel ** 2).sum()),
            "internal_energy": float((mass * e).sum()),
            "mode1": float((vel[:, 0] * np.sin(wavenumber * pos[:, 2])).sum()),
            "mode2": float((vel[:, 0] * np.sin(wavenumber * pos[:, 1])
                            * np.sin(wavenumber * pos[:, 2])).sum()),
            "absmv": float((m * np.abs(vel)).sum())}


def shear_wave_loop(pos, vel, mass, *, h, box, rho0, sound_speed, gamma=7.0, nu, dt, steps,
                    skin, renormalise=True, sample_every=10, wavenumber=0.0, viscous_order=2):
    """The step loop of threed.solver.run_shear_wave (solver.py:254-308), FP64:
    periodic pairs at 2h + skin, rebuilt when 2 max|min_image(x - x_anchor)| >
    skin; kick-drift-kick with x wrapped by the floored remainder."""
    pos = np.array(pos, np.float64); vel = np.array(vel, np.float64)
    mass = np.asarray(mass, np.float64)
    cutoff = 2.0 * h + skin
    pairs = periodic_pairs(pos, box, cutoff)
    anchor = pos.copy()
    rebuilds = 1
    kw = dict(rho0=rho0, sound_speed=sound_speed, gamma=gamma, viscosity=nu, box=box,
              renormalise=renormalise, viscous_order=viscous_order)
    acc, rho, _ = threed_accelerations(pos, vel, mass, h, pairs, **kw)
    rows = []
    for k in range(steps + 1):
        if k % sample_every == 0:
            rows.append(ledger_row(pos, vel, mass, rho, rho0=rho0, sound_speed=sound_speed,
                                   gamma=gamma, box=box, wavenumber=wavenumber))
        if k == steps:
            break
        vel = vel + 0.5 * dt * acc
        pos = np.remainder(pos + dt * vel, box)
        d = pos - anchor
        d = d - box * np.round(d / box)
        if 2.0 * float(np.sqrt((d * d).sum(1)).max()) > skin:
            pairs = periodic_pairs(pos, box, cutoff)
            anchor = pos.copy()
            rebuilds += 1
        acc, rho, _ = threed_accelerations(pos, vel, mass, h, pairs, **kw)
        vel = vel + 0.5 * dt * acc
    return {"pos": pos, "vel": vel, "rho": rho, "acc": acc, "rows": rows, "rebuilds": rebuilds,
            "pairs_last": int(pairs.shape[0])}


def clamp_out_of_obstacle(pos, vel, mass, lo, hi, eps):
    """kleefsman_obstacle_3d.clamp_out_of_obstacle (kleefsman_obstacle_3d.py:208-266):
    particles strictly inside the box leave through the nearest face (first of
    -x, +x, -y, +y, -z, +z on a tie) to face -/+ eps, that velocity component
    is zeroed.  Returns (pos, vel, count, impulse = sum m (v_before - v_after))."""
    pos = np.array(pos, np.float64); vel = np.array(vel, np.float64)
    lo = np.asarray(lo, np.float64); hi = np.asarray(hi, np.float64)
    inside = np.all((pos > lo) & (pos < hi), axis=1)
    impulse = np.zeros(3)
    idx = np.nonzero(inside)[0]
    for i in idx:
        cand = [pos[i, 0] - lo[0], hi[0] - pos[i, 0], pos[i, 1] - lo[1], hi[1] - pos[i, 1],
                pos[i, 2] - lo[2], hi[2] - pos[i, 2]]
        face = int(np.argmin(cand))
        d = face // 2
        pos[i, d] = hi[d] + eps if face % 2 else lo[d] - eps
        impulse[d] += mass[i] * vel[i, d]
        vel[i, d] = 0.0
    return pos, vel, int(idx.size), impulse


def apply_boundary(pos, vel, lo, hi, tol):
    pos = np.ascontiguousarray(pos, np.float64); vel = np.ascontiguousarray(vel, np.float64)
    dim = pos.shape[1]
    lo = np.ascontiguousarray(lo, np.float64); hi = np.ascontiguousarray(hi, np.float64)
    cnt = lib().orc_apply_boundary(dim, pos.shape[0], _d(pos), _d(vel), _d(lo), _d(hi), tol)
    return pos, vel, int(cnt)


def delta_plus_step(pos, vel, rho, mass, *, dx, sound_speed, reference_speed, dt, rho0=1000.0,
                    gravity=9.81, viscosity=1e-6, diffusion=0.1, alpha=0.02, wall=None,
                    wall_velocity=None, wall_normal=None, pairs=None):
    """delta_plus.delta_plus_step (delta_plus.py:105-229), FP64.

    Returns dict(position, velocity, density, acceleration, wall_reaction, shift)."""
    pos = np.ascontiguousarray(pos, np.float64); vel = np.ascontiguousarray(vel, np.float64)
    rho = np.ascontiguousarray(rho, np.float64); mass = np.ascontiguousarray(mass, np.float64)
    N = pos.shape[0]
    wall = np.zeros((0, 2)) if wall is None else np.ascontiguousarray(wall, np.float64)
    W = wall.shape[0]
    wv = np.zeros((W, 2)) if wall_velocity is None else np.ascontiguousarray(wall_velocity, np.float64)
    wn = np.zeros((W, 2)) if wall_normal is None else np.asarray(wall_normal, np.float64)
    if W:
        wn = np.ascontiguousarray(wn / np.linalg.norm(wn, axis=1)[:, None])
    if pairs is None:
        pairs = cell_linked_pairs(np.concatenate([pos, wall]), 4 * dx)
    pairs = np.ascontiguousarray(pairs, np.int64).reshape(-1, 2)
    P = DPParams(dx, sound_speed, reference_speed, dt, rho0, gravity, viscosity, diffusion, alpha)
    out = {k: np.zeros((N, 2)) for k in ("position", "velocity", "acceleration", "shift")}
    out["density"] = np.zeros(N)
    out["wall_reaction"] = np.zeros((W, 2))
    st = lib().orc_delta_plus_step(C.byref(P), N, W, _d(pos), _d(vel), _d(rho), _d(mass),
                                   _d(wall), _d(wv), _d(wn), _l(pairs), pairs.shape[0],
                                   _d(out["position"]), _d(out["velocity"]), _d(out["density"]),
                                   _d(out["acceleration"]), _d(out["wall_reaction"]),
                                   _d(out["shift"]))
    if st == 1:
        raise FloatingPointError("nonpositive extrapolated wall density")
    if st == 2:
        raise FloatingPointError("delta-plus density became nonpositive or nonfinite")
    return out


# --------------------------------------------------------------------------
# Step drivers (2-D WCSPH loop of droplet_impact.run :1674-1763), used as the
# CPU baseline and for trajectory parity.
# --------------------------------------------------------------------------

def lattice_kernel_sum_2d(h, dx):
    """droplet_impact.lattice_kernel_sum (droplet_impact.py:577-598)."""
    reach = int(math.ceil(2.0 * h / dx)) + 1
    off = np.arange(-reach, reach + 1, dtype=np.float64) * dx
    gx, gy = np.meshgrid(off, off, indexing="xy")
    r = np.sqrt(gx ** 2 + gy ** 2).reshape(-1)
    w, _, _ = kernel_eval(0, r, h)
    return float(w.sum())


class Stepper3D:
    """The step loop of kleefsman_obstacle_3d.run_case (:553-726) in FP64:
    Verlet pair cache with wall-wall pairs pruned, rhs_3d with wall reactions,
    obstacle force / moment, symplectic Euler, apply_boundary_3d and
    clamp_out_of_obstacle."""

    def __init__(self, pos, vel, mass, wall, *, h, rho0, c0, gamma, gravity, nu, dt, wall_mass,
                 skin, lo, hi, tol, obstacle_rows, centre, box=None, eps=0.0, wall_friction=0.0):
        self.pos = np.array(pos, np.float64); self.vel = np.array(vel, np.float64)
        self.mass = np.asarray(mass, np.float64); self.wall = np.asarray(wall, np.float64)
        self.kw = dict(dim=3, h=h, rho0=rho0, c0=c0, gamma=gamma, gravity=gravity, nu=nu,
                       wall_friction=wall_friction, dt=dt, wall_mass=wall_mass, kernel3d=True,
                       reaction_style=2)
        self.h, self.dt, self.skin = h, dt, skin
        self.lo, self.hi, self.tol = lo, hi, tol
        self.rows = slice(*obstacle_rows)
        self.centre = np.asarray(centre, np.float64)
        self.box, self.eps = box, eps
        self.anchor = None
        self.pairs = None
        self.rebuilds = 0
        self.clamped = 0
        self.clamp_impulse = np.zeros(3)
        self.loads = []

    def _pairs(self):
        n = self.pos.shape[0]
        allp = np.concatenate([self.pos, self.wall])
        disp = math.inf if self.anchor is None else float(np.sqrt(((allp - self.anchor) ** 2).sum(1)).max())
        if self.pairs is None or disp > 0.5 * self.skin:
            pairs = cell_linked_pairs(allp, 2.0 * self.h + self.skin)
            keep = (pairs[:, 0] < n) | (pairs[:, 1] < n)
            self.pairs = pairs[keep]
            self.anchor = allp.copy()
            self.rebuilds += 1
        return self.pairs

    def step(self):
        pairs = self._pairs()
        rho, p, acc, _, _, reac = rhs_morris(self.pos, self.vel, self.mass, self.wall, pairs, **self.kw)
        loads = reac[self.rows]
        force = loads.sum(0)
        moment = np.cross(self.wall[self.rows] - self.centre, loads).sum(0)
        self.loads.append(np.concatenate([force, moment]))
        self.vel = self.vel + self.dt * acc
        self.pos = self.pos + self.dt * self.vel
        self.pos, self.vel, _ = apply_boundary(self.pos, self.vel, self.lo, self.hi, self.tol)
        if self.box is not None:
            self.pos, self.vel, cnt, imp = clamp_out_of_obstacle(self.pos, self.vel, self.mass,
                                                                 self.box[0], self.box[1], self.eps)
            self.clamped += cnt
            self.clamp_impulse += imp
        return rho, p, acc


class Stepper2D:
    """Symplectic-Euler WCSPH loop with the Verlet pair cache (FP64)."""

    def __init__(self, pos, vel, mass, wall, *, h, rho0, c0, gamma, gravity, nu, dt,
                 wall_mass, skin, lo, hi, tol, wall_friction=0.0, diff_scale=0.0,
                 diff_form=0, eig_floor=0.1, **options):
        """``options``: the rhs_morris opt-in branches (shepard, lattice_sum,
        bcd, closure, closure_threshold, closure_literal, closure_layers)."""
        self.pos = np.array(pos, np.float64); self.vel = np.array(vel, np.float64)
        self.mass = np.asarray(mass, np.float64); self.wall = np.asarray(wall, np.float64)
        self.kw = dict(dim=2, h=h, rho0=rho0, c0=c0, gamma=gamma, gravity=gravity, nu=nu,
                       wall_friction=wall_friction, dt=dt, diff_scale=diff_scale,
                       diff_form=diff_form, eig_floor=eig_floor, wall_mass=wall_mass, **options)
        self.h, self.dt, self.skin = h, dt, skin
        self.lo, self.hi, self.tol = lo, hi, tol
        self.anchor = None
        self.pairs = None
        self.rebuilds = 0
        self.escaped = 0

    def pairs_for_step(self):
        n = self.pos.shape[0]
        allp = np.concatenate([self.pos, self.wall])
        if self.anchor is not None:
            disp = float(np.sqrt(((allp - self.anchor) ** 2).sum(1)).max()) if len(allp) else 0.0
        else:
            disp = math.inf
        if self.pairs is None or disp > 0.5 * self.skin:
            pairs = cell_linked_pairs(allp, 2.0 * self.h + self.skin)
            keep = (pairs[:, 0] < n) | (pairs[:, 1] < n)
            self.pairs = pairs[keep]
            self.anchor = allp.copy()
            self.rebuilds += 1
        return self.pairs

    def step(self):
        pairs = self.pairs_for_step()
        rho, p, acc, _, _, _ = rhs_morris(self.pos, self.vel, self.mass, self.wall, pairs,
                                          **self.kw)
        self.vel = self.vel + self.dt * acc
        self.pos = self.pos + self.dt * self.vel
        self.pos, self.vel, cnt = apply_boundary(self.pos, self.vel, self.lo, self.hi, self.tol)
        self.escaped += cnt
        return rho, p, acc


def adaptive_next_dt(dt: float, max_speed: float, *, dt_cfg: float, max_courant: float, h_factor: float,
                     dx: float, c0: float, viscosity: float, gravity: float, diffusion: float,
                     dt_safety: float, dt_growth: float) -> float:
    """The adaptive step after one step (droplet_impact.py:495-552, :1389-1421):
    the smallest of the acoustic limit at the step's starting speed
    max_courant h / (c0 + |v|), 0.125 h^2 / nu, 0.25 sqrt(h / g) and
    0.125 h / (delta c0), scaled by dt_safety; grown from dt by at most
    dt_growth (or taken at once when dt_growth < 1); never above dt_cfg."""
    h = h_factor * dx
    acoustic = max_courant * h_factor * dx / (c0 + max(0.0, max_speed))
    viscous = 0.125 * h * h / viscosity if viscosity > 0.0 else math.inf
    grav = 0.25 * math.sqrt(h_factor * dx / gravity) if gravity > 0.0 else math.inf
    diff = 0.125 * h / (diffusion * c0) if diffusion > 0.0 else math.inf
    target = dt_safety * min(acoustic, viscous, grav, diff)
    grown = dt * dt_growth if dt_growth >= 1.0 else target
    return min(target, grown, dt_cfg)


def _probe_cubic(q):
    q = np.asarray(q, np.float64)
    return np.where(q < 1.0, 1.0 - 1.5 * q * q + 0.75 * q ** 3,
                    0.25 * np.clip(2.0 - q, 0.0, None) ** 3)


def kleefsman_gauges(pos, volume, rho, pressure, *, station_x, aperture, tank_depth, dx, h, gravity,
                     obstacle_lo, obstacle_hi, nodes):
    """The gauges of kleefsman_obstacle_3d.py:273-318 and marin_probes.py:23-69
    in FP64: slab height n dx^3 / (aperture * box width) over the band
    |x - station_x| < aperture / 2, box z-span; the highest y in the band (0
    if empty); the front max x; the bypass count (x past the box, z outside
    its span); the centreline height sum V cubic(r / h) * 10 / (7 pi h^2),
    r = hypot(x - station_x, z - depth / 2); and per quadrature node (taps 0-3
    upstream face: x <= x_node, taps 4-7 top face: y >= y_node) the pressure
    max(0, sum w (p + rho g (y - y_node)) / max(sum w, tiny)) with
    w = cubic(|x - node| / h) V, and the count of w > 0."""
    pos = np.asarray(pos, np.float64)
    vol = np.asarray(volume, np.float64)
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    oz0, oz1 = obstacle_lo[2], obstacle_hi[2]
    band = (x > station_x - 0.5 * aperture) & (x < station_x + 0.5 * aperture) & (z > oz0) & (z < oz1)
    n = int(band.sum())
    slab = n * dx ** 3 / (aperture * (oz1 - oz0))
    peak = float(y[band].max()) if n else 0.0
    front = float(x.max()) if len(x) else 0.0
    bypass = int(((x > obstacle_hi[0]) & ((z < oz0) | (z > oz1))).sum())
    r = np.hypot(x - station_x, z - 0.5 * tank_depth)
    h2 = float((vol * _probe_cubic(r / h)).sum() * (10.0 / (7.0 * np.pi * h * h)))
    nodes = np.asarray(nodes, np.float64)
    delta = pos[:, None, None, :] - nodes[None]
    w = _probe_cubic(np.linalg.norm(delta, axis=-1) / h) * vol[:, None, None]
    w[:, :4] *= delta[:, :4, :, 0] <= 0
    w[:, 4:] *= delta[:, 4:, :, 1] >= 0
    corr = (np.asarray(pressure, np.float64)[:, None, None]
            + np.asarray(rho, np.float64)[:, None, None] * gravity * delta[..., 1])
    den = w.sum(0)
    p = np.clip((w * corr).sum(0) / np.maximum(den, np.finfo(np.float64).tiny), 0.0, None)
    return {"slab": slab, "peak": peak, "front": front, "bypass": bypass, "h2": h2,
            "tap_p": p, "tap_support": (w > 0).sum(0)}


def lind_shift(pos, volume, h, *, dt, v_ref, strength=1.0, limit=0.25, free_surface=0.95,
               percentile=95.0, c_reference=None):
    """particle_shifting.shift_positions (particle_shifting.py:88-224), 2-D,
    FP64: C_i = V_i W(0) + sum_j V_j W_ij over pairs within 2h; grad C_i =
    sum_j V_j (C_j - C_i) grad_i W_ij; shift = -strength h v_ref dt grad C;
    rows with C_i < free_surface * C_ref (C_ref: the percentile of C, linear
    interpolation) keep only the part of the shift normal to n = -grad C /
    |grad C|; |shift| is capped at limit * h.  Returns (new positions, C,
    C_ref, free-surface mask)."""
    pos = np.asarray(pos, np.float64)
    vol = np.asarray(volume, np.float64)
    pairs = cell_linked_pairs(pos, 2.0 * h)
    i, j = pairs[:, 0], pairs[:, 1]
    d = pos[i] - pos[j]
    r = np.sqrt((d * d).sum(1))
    w0, _, _ = kernel_eval(0, np.zeros(1), h)
    w, g, _ = kernel_eval(0, r, h)          # 2-D cubic spline; g = dW/dr / r
    c = vol * w0[0]
    np.add.at(c, i, vol[j] * w)
    np.add.at(c, j, vol[i] * w)
    gw = d * g[:, None]
    diff = (c[j] - c[i])[:, None] * gw
    grad = np.zeros_like(pos)
    np.add.at(grad, i, vol[j, None] * diff)
    np.add.at(grad, j, vol[i, None] * diff)
    norm = np.sqrt((grad * grad).sum(1))
    normal = np.where(norm[:, None] > 1e-30, -grad / np.maximum(norm, 1e-30)[:, None], 0.0)
    c_ref = float(np.percentile(c, percentile)) if c_reference is None else float(c_reference)
    fs = c < free_surface * c_ref if c_ref > 0 else np.zeros(len(c), bool)
    shift = -(strength * h * v_ref * dt) * grad
    proj = (shift * normal).sum(1, keepdims=True) * normal
    shift = np.where(fs[:, None], shift - proj, shift)
    cap = limit * h
    mag = np.sqrt((shift * shift).sum(1))
    shift = np.where((mag > cap)[:, None], shift * (cap / np.maximum(mag, 1e-300))[:, None], shift)
    return pos + shift, c, c_ref, fs


def tvc_shift(pos, volume, h, n_fluid, pairs, *, alpha=0.05, slope=None, max_shift=None, active=None,
              scope_tol=None):
    """transport_velocity_correction (transport_velocity_correction.py:285-513)
    in 2-D FP64 with unit kernel correction: g_i = -sum_j f_ij V_j grad_i W_ij
    over the pairs (f = 2 for fluid-fluid, the wall factor 2 for fluid-wall),
    shift = alpha h^2 L(h^2 |g|^2) g with L = 1 or min(slope x, 1), zeroed
    outside ``active`` and where |g| h > scope_tol, capped at max_shift.
    Returns (g, shift) over the fluid rows."""
    pos = np.asarray(pos, np.float64)
    vol = np.asarray(volume, np.float64)
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    i, j = pairs[:, 0], pairs[:, 1]
    d = pos[i] - pos[j]
    r = np.sqrt((d * d).sum(1))
    _, gs, _ = kernel_eval(0, r, h)
    gw = d * gs[:, None]                       # grad_i W_ij; grad_j W_ij = -gw
    g = np.zeros((n_fluid, 2))
    fi, fj = i < n_fluid, j < n_fluid
    ff = fi & fj
    np.add.at(g, i[ff], (2.0 * vol[j[ff]])[:, None] * gw[ff])
    np.add.at(g, j[ff], (2.0 * vol[i[ff]])[:, None] * -gw[ff])
    fw = fi & ~fj
    np.add.at(g, i[fw], (2.0 * vol[j[fw]])[:, None] * gw[fw])
    wf = ~fi & fj
    np.add.at(g, j[wf], (2.0 * vol[i[wf]])[:, None] * -gw[wf])
    g = -g
    sq = (g * g).sum(1)
    fac = np.ones_like(sq) if slope is None else np.minimum(slope * h * h * sq, 1.0)
    shift = (alpha * h * h) * fac[:, None] * g
    if active is not None:
        shift = shift * np.asarray(active, np.float64)[:, None]
    if scope_tol is not None:
        shift = shift * (np.sqrt(sq) * h <= scope_tol)[:, None]
    if max_shift is not None:
        nrm = np.sqrt((shift * shift).sum(1))
        over = nrm > max_shift
        shift = np.where(over[:, None], shift * (max_shift / np.maximum(nrm, 1e-30))[:, None], shift)
    return g, shift


# ---------------------------------------------------------------------------
# Multi-threaded CPU port of the 2-D dam-break step (tsph_cpu_omp.c): the
# benchmark's CPU baseline, the same step as Stepper2D (default rhs branch)
# formulated per row so that it runs on every host core.

OMP_LIB_PATH = HERE / "libtsph_cpu_omp.so"
_omp = None


class OmpParams(C.Structure):
    _fields_ = [("h", C.c_double), ("rho0", C.c_double), ("c0", C.c_double), ("gamma", C.c_double),
                ("gravity", C.c_double), ("nu", C.c_double), ("dt", C.c_double), ("skin", C.c_double),
                ("wall_mass", C.c_double), ("lo", C.c_double * 2), ("hi", C.c_double * 2),
                ("tol", C.c_double)]


def omp_lib():
    global _omp
    if _omp is None:
        if not OMP_LIB_PATH.exists():
            build()
        _omp = C.CDLL(str(OMP_LIB_PATH))
        _omp.omp_sim_create.restype = C.c_void_p
        _omp.omp_sim_create.argtypes = [C.POINTER(OmpParams), C.c_int64, C.c_int64, _dp, _dp, _dp, _dp]
        _omp.omp_sim_step.argtypes = [C.c_void_p, C.c_int64]
        _omp.omp_sim_read.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _lp]
        _omp.omp_sim_destroy.argtypes = [C.c_void_p]
        _omp.omp_threads.argtypes = [C.c_int]
        _omp.omp_threads.restype = C.c_int
    return _omp


class OmpStepper2D:
    """Stepper2D's step (no opt-in branches) on ``threads`` host threads."""

    def __init__(self, pos, vel, mass, wall, *, h, rho0, c0, gamma, gravity, nu, dt, wall_mass, skin,
                 lo, hi, tol, threads=0):
        self.threads = omp_lib().omp_threads(int(threads))
        pos = np.ascontiguousarray(pos, np.float64); vel = np.ascontiguousarray(vel, np.float64)
        mass = np.ascontiguousarray(mass, np.float64); wall = np.ascontiguousarray(wall, np.float64)
        self.n = pos.shape[0]
        P = OmpParams(h, rho0, c0, gamma, gravity, nu, dt, skin, wall_mass,
                      (C.c_double * 2)(*lo), (C.c_double * 2)(*hi), tol)
        self._h = omp_lib().omp_sim_create(C.byref(P), self.n, wall.shape[0], _d(pos), _d(vel), _d(mass),
                                           _d(wall))

    def step(self, k=1):
        omp_lib().omp_sim_step(self._h, int(k))

    def read(self):
        pos = np.zeros((self.n, 2)); vel = np.zeros((self.n, 2)); rho = np.zeros(self.n)
        acc = np.zeros((self.n, 2)); rb = C.c_int64(0)
        omp_lib().omp_sim_read(self._h, _d(pos), _d(vel), _d(rho), _d(acc), C.byref(rb))
        return {"position": pos, "velocity": vel, "density": rho, "acceleration": acc,
                "rebuilds": rb.value}

    def close(self):
        if self._h:
            omp_lib().omp_sim_destroy(self._h)
            self._h = None
