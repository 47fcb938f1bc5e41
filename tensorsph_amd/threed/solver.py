"""3-D drivers of tensorsph.threed.solver (solver.py:65-341) on the resident
periodic stepper of libtsph (tsph_psim_*).

``run_shear_wave`` keeps the reference's signature and result object; the
loop runs on the GPU: the pair list (periodic cells at 2h + skin, rebuilt when
the largest minimum-image displacement exceeds skin / 2), ``accelerations``,
the kick-drift-kick leapfrog with positions wrapped into the box, and the
conservation ledger rows all stay in HBM; only the ledger rows come back.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _lib
from .conservation import ROW, R_ABSMV, R_MODE1, R_MODE2, row_to_ledger
from .kernel3 import calibrated_particle_mass
from .operators3 import _FORMS, _params, analytic_mode_decay

COURANT = 0.25


def acoustic_dt(dx: float, h: float, sound_speed: float, courant: float = COURANT,
                sigma: float = 0.0, rho0: float = 1000.0) -> float:
    """courant h / c0, capped by the grid-scale capillary limit
    courant sqrt(rho0 h^3 / (sigma pi^3)) when sigma > 0 (solver.py:65-81)."""
    dt = courant * h / sound_speed
    if sigma > 0.0:
        dt = min(dt, courant * math.sqrt(rho0 * h ** 3 / (sigma * math.pi ** 3)))
    return dt


def cubic_lattice(shape, dx: float, origin=(0.0, 0.0, 0.0), dtype=torch.float64) -> torch.Tensor:
    """Points i dx + origin on an (nx, ny, nz) lattice, last axis fastest
    (solver.py:84-90).  Host tensor, like the reference's."""
    axes = [torch.arange(int(n), dtype=dtype) * dx + origin[a] for a, n in enumerate(shape)]
    return torch.stack(torch.meshgrid(*axes, indexing="ij"), dim=-1).reshape(-1, len(shape))


class PeriodicSimulation:
    """A triply periodic fluid advanced by libtsph's resident leapfrog stepper."""

    def __init__(self, pos, vel, mass, *, h: float, box, rho0: float, sound_speed: float,
                 gamma: float = 7.0, viscosity: float = 0.0, dt: float, skin: float,
                 renormalise: bool = True, density_diffusion: float = 0.0,
                 density_diffusion_form: str = "antuono", eigenvalue_floor: float = 0.1,
                 precision: str = "fp64", device=None, viscous_order: int = 2):
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if dev.type != "cuda":
            raise ValueError("PeriodicSimulation runs on a ROCm device")
        self.device = dev
        self.code = _lib.F64 if precision == "fp64" else _lib.F32
        self.sdt = torch.float64 if precision == "fp64" else torch.float32
        x = torch.as_tensor(np.asarray(pos, dtype=np.float64) if not isinstance(pos, torch.Tensor) else pos)
        v = torch.as_tensor(np.asarray(vel, dtype=np.float64) if not isinstance(vel, torch.Tensor) else vel)
        m = torch.as_tensor(np.asarray(mass, dtype=np.float64) if not isinstance(mass, torch.Tensor) else mass)
        self.n = int(x.shape[0])
        if x.shape != (self.n, 3) or v.shape != (self.n, 3) or m.shape != (self.n,):
            raise ValueError("pos, vel must be (N, 3) and mass (N,)")
        self._x = x.to(dev, torch.float64).contiguous()
        self._v = v.to(dev, torch.float64).contiguous()
        self._m = m.to(dev, self.sdt).contiguous()
        self.box = [float(b) for b in torch.as_tensor(box, dtype=torch.float64).reshape(-1).tolist()]
        if len(self.box) == 1:
            self.box = self.box * 3
        if density_diffusion > 0.0 and density_diffusion_form not in _FORMS:
            raise ValueError(f"unknown density_diffusion_form {density_diffusion_form!r}")
        self.h, self.rho0, self.c0, self.gamma = float(h), float(rho0), float(sound_speed), float(gamma)
        self.dt, self.skin = float(dt), float(skin)
        self.prm = _params(self.code, self.n, self.n, h, rho0=rho0, sound_speed=sound_speed, gamma=gamma,
                           viscosity=viscosity, box=self.box, renormalise=renormalise,
                           diffusion_scale=(2.0 * density_diffusion * sound_speed * h
                                            if density_diffusion > 0.0 else 0.0),
                           diffusion_form=_FORMS.get(density_diffusion_form, 1),
                           eigenvalue_floor=eigenvalue_floor, dt=dt, viscous_order=viscous_order)
        handle = C.c_void_p()
        with torch.cuda.device(dev):
            _lib.check(_lib.lib().tsph_psim_create(C.byref(self.prm), self.skin, _lib.ptr(self._x),
                                                   _lib.ptr(self._v), _lib.ptr(self._m),
                                                   _lib.stream_handle(dev), C.byref(handle)))
        self._h = handle
        if self.code == _lib.F32:
            # FP32 steps keep each mass as two floats and sum densities with
            # the FP64 value (near rho0 a mass rounding shifts p by ~7 B 2^-24)
            self._m64 = m.to(dev, torch.float64).contiguous()
            with torch.cuda.device(dev):
                _lib.check(_lib.lib().tsph_psim_set_masses_f64(handle, _lib.ptr(self._m64),
                                                               _lib.stream_handle(dev)))

    def _stream(self):
        return _lib.stream_handle(self.device)

    def step(self, nsteps: int = 1) -> None:
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().tsph_psim_step(self._h, int(nsteps), self._stream()))

    def row(self, wavenumber: float = 0.0) -> list:
        out = (C.c_double * ROW)()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().tsph_psim_diagnostics(self._h, float(wavenumber), out, self._stream()))
        return list(out)

    def diagnostics(self) -> dict:
        """conservation.diagnostics of the current state (periodic centre of mass)."""
        return row_to_ledger(self.row(), self.box)

    def read(self) -> dict:
        pos = torch.empty((self.n, 3), dtype=torch.float64, device=self.device)
        vel = torch.empty_like(pos)
        rho = torch.empty(self.n, dtype=self.sdt, device=self.device)
        acc = torch.empty((self.n, 3), dtype=self.sdt, device=self.device)
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().tsph_psim_read(self._h, _lib.ptr(pos), _lib.ptr(vel), _lib.ptr(rho),
                                                 _lib.ptr(acc), self._stream()))
        torch.cuda.synchronize(self.device)
        return {"position": pos, "velocity": vel, "density": rho, "acceleration": acc}

    def stats(self) -> dict:
        c = (C.c_int64 * 3)()
        _lib.check(_lib.lib().tsph_psim_stats(self._h, c))
        return {"steps": c[0], "rebuilds": c[1], "list_entries": c[2], "pairs": c[2] // 2}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().tsph_psim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class ShearWaveResult:
    """Outcome of one shear-wave run (solver.py:134-208)."""

    n: int
    dx: float
    h: float
    nu: float
    wavenumber: float
    sound_speed: float
    amplitude0: float
    dt: float
    steps: int
    renormalise: bool
    viscous_order: int = 2
    times: list = field(default_factory=list)
    amplitudes: list = field(default_factory=list)
    fit_rate: float = float("nan")
    analytic_rate: float = float("nan")
    diagnostics: list = field(default_factory=list)
    pairs_last: int = 0
    pair_rebuilds: int = 0
    momentum_scale: float = 0.0

    @property
    def rate_error(self) -> float:
        return (self.fit_rate - self.analytic_rate) / self.analytic_rate

    def as_dict(self) -> dict:
        first, last = self.diagnostics[0], self.diagnostics[-1]
        scale = self.momentum_scale
        return {
            "case": "shear_wave_3d", "n": self.n, "particles": self.n ** 3, "dx": self.dx,
            "h": self.h, "h_over_dx": self.h / self.dx, "kh": self.wavenumber * self.h,
            "nu": self.nu, "wavenumber": self.wavenumber, "sound_speed": self.sound_speed,
            "amplitude0": self.amplitude0, "mach": self.amplitude0 / self.sound_speed,
            "dt": self.dt, "steps": self.steps, "renormalise_viscous_operator": self.renormalise,
            "viscous_operator_order": self.viscous_order,
            "decay": {"fit_rate": self.fit_rate, "analytic_rate": self.analytic_rate,
                      "rel_error": self.rate_error,
                      "amplitude_final": self.amplitudes[-1] if self.amplitudes else None},
            "conservation": {
                "momentum_scale_sum_m_abs_v": scale,
                "momentum_final_over_scale": (last["momentum_magnitude"] / scale
                                              if scale > 0.0 else None),
                "mass_relative_drift": (last["mass"] - first["mass"]) / first["mass"],
                "momentum_initial": first["momentum"], "momentum_final": last["momentum"],
                "momentum_magnitude_initial": first["momentum_magnitude"],
                "momentum_magnitude_final": last["momentum_magnitude"],
                "momentum_magnitude_peak": max(d["momentum_magnitude"] for d in self.diagnostics),
                "kinetic_energy_initial": first["kinetic_energy"],
                "kinetic_energy_final": last["kinetic_energy"],
                "total_energy_initial": first["total_energy"],
                "total_energy_final": last["total_energy"],
                "total_energy_relative_drift": ((last["total_energy"] - first["total_energy"])
                                                / first["total_energy"]),
                "angular_momentum_drift": [b - a for a, b in zip(first["angular_momentum"],
                                                                 last["angular_momentum"])],
            },
            "pairs_last": self.pairs_last, "pair_rebuilds": self.pair_rebuilds,
            "series": {"times": self.times, "amplitudes": self.amplitudes},
        }


def fit_decay_rate(times, amplitudes) -> float:
    """Least-squares slope of -ln|a| against t, first sample dropped (solver.py:311-327)."""
    t = np.asarray(times[1:], dtype=np.float64)
    a = np.abs(np.asarray([float(x) for x in amplitudes[1:]], dtype=np.float64))
    keep = a > 0.0
    t, a = t[keep], a[keep]
    if t.size < 2:
        return float("nan")
    tc = t - t.mean()
    la = np.log(a)
    return float(-(tc * (la - la.mean())).sum() / (tc * tc).sum())


def shear_wave_velocity(pos: torch.Tensor, amplitude: float, wavenumber: float,
                        mode: str = "double") -> torch.Tensor:
    """u_x = A sin(k y) sin(k z) ("double", operators3.py:416-421) or
    u_x = A sin(k z) ("single")."""
    vel = torch.zeros_like(pos)
    if mode == "double":
        vel[:, 0] = amplitude * torch.sin(wavenumber * pos[:, 1]) * torch.sin(wavenumber * pos[:, 2])
    elif mode == "single":
        vel[:, 0] = amplitude * torch.sin(wavenumber * pos[:, 2])
    else:
        raise ValueError("mode must be 'double' or 'single'")
    return vel


def run_shear_wave(*, n: int = 16, box: float = 1.0, h_factor: float = 1.3, periods: int = 1,
                   nu: float = 0.02, amplitude: float = 0.1, sound_speed: float = 10.0,
                   rho0: float = 1000.0, gamma: float = 7.0, t_end: float | None = None,
                   dt: float | None = None, steps: int | None = None, sample_every: int = 10,
                   skin_factor: float = 0.25, renormalise: bool = True, viscous_order: int = 2,
                   dtype=torch.float64, device=None, mode: str = "double",
                   jitter: float = 0.0, seed: int = 20261203) -> ShearWaveResult:
    """Integrate the 3-D shear field and fit its decay exponent (solver.py:211-308).

    Extra keyword arguments beyond the reference's: ``device``; ``mode``
    ("single" gives u_x = A sin(k z), decay exp(-nu k^2 t)); ``jitter``, a
    uniform perturbation of +-jitter dx per coordinate drawn from numpy's
    PCG64(seed).  dtype float64 runs the FP64 kernels, float32 the
    FP32-compensated ones (positions stay FP64 on the device).
    """
    dx = box / n
    h = h_factor * dx
    wavenumber = 2.0 * math.pi * periods / box
    rate = analytic_mode_decay(nu, wavenumber) if mode == "double" else nu * wavenumber ** 2
    if t_end is None:
        t_end = 2.0 / rate
    if dt is None:
        dt = acoustic_dt(dx, h, sound_speed)
    if steps is None:
        steps = int(math.ceil(t_end / dt))
    viscous_limit = 0.125 * h * h / max(nu, 1.0e-30)
    if dt > viscous_limit:
        raise ValueError(f"dt = {dt:.3e} s exceeds the viscous limit 0.125 h^2/nu = "
                         f"{viscous_limit:.3e} s")
    if wavenumber * h > 0.9:
        raise ValueError(f"k h = {wavenumber * h:.2f}: the kernel cannot represent this mode")
    pos = cubic_lattice((n, n, n), dx, dtype=dtype)
    if jitter > 0.0:
        rng = np.random.Generator(np.random.PCG64(seed))
        pos = (pos.double() + torch.from_numpy(rng.uniform(-jitter * dx, jitter * dx, size=tuple(pos.shape))))
        pos = pos.remainder(box).to(dtype)
    mass = torch.full((pos.shape[0],), calibrated_particle_mass(dx, h, rho0), dtype=dtype)
    vel = shear_wave_velocity(pos, amplitude, wavenumber, mode)
    cutoff = 2.0 * h * (1.0 + skin_factor)
    skin = cutoff - 2.0 * h
    sim = PeriodicSimulation(pos, vel, mass, h=h, box=box, rho0=rho0, sound_speed=sound_speed,
                             gamma=gamma, viscosity=nu, dt=dt, skin=skin, renormalise=renormalise,
                             precision="fp64" if dtype == torch.float64 else "fp32", device=device,
                             viscous_order=viscous_order)
    try:
        result = ShearWaveResult(n=n, dx=dx, h=h, nu=nu, wavenumber=wavenumber,
                                 sound_speed=sound_speed, amplitude0=amplitude, dt=dt, steps=steps,
                                 renormalise=renormalise, viscous_order=viscous_order)
        npart = pos.shape[0]
        proj, weight = (R_MODE2, 4.0) if mode == "double" else (R_MODE1, 2.0)
        k = 0
        while True:
            if k % sample_every == 0:
                row = sim.row(wavenumber)
                if k == 0:
                    result.momentum_scale = row[R_ABSMV]
                result.times.append(k * dt)
                result.amplitudes.append(weight * row[proj] / npart)
                result.diagnostics.append(row_to_ledger(row, [box] * 3))
            if k == steps:
                break
            m = min(sample_every - k % sample_every, steps - k)
            sim.step(m)
            k += m
        st = sim.stats()
        result.pairs_last = st["pairs"]
        result.pair_rebuilds = st["rebuilds"]
        result.analytic_rate = rate
        result.fit_rate = fit_decay_rate(result.times, result.amplitudes)
        return result
    finally:
        sim.close()
