"""Conservation ledger of tensorsph.threed.conservation (conservation.py:41-145).

``diagnostics`` reduces the whole row in one HIP pass (tsph_conservation_row,
FP64 partial sums in a fixed order); the periodic stepper uses the same
reduction on its resident state.  The single-quantity helpers are thin
expressions over the caller's device tensors.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from .. import _lib

ROW = 21
(R_MASS, R_P, R_L, R_KE, R_IE, R_SIN, R_COS, R_MODE1, R_MODE2, R_ABSMV, R_MX) = (
    0, 1, 4, 7, 8, 9, 12, 15, 16, 17, 18)


def conservation_row(pos, vel, mass, rho, *, rho0: float, sound_speed: float, gamma: float = 7.0,
                     box=None, wavenumber: float = 0.0) -> list:
    """The 21-value FP64 ledger row of libtsph for (n, 3) pos / vel."""
    dev = _lib.require_gpu(pos, vel, mass, rho)
    if pos.ndim != 2 or pos.shape[1] != 3 or vel.shape != pos.shape:
        raise ValueError("pos and vel must have shape (N, 3)")
    code = _lib.F32 if mass.dtype == torch.float32 else _lib.F64
    sdt = torch.float32 if code == _lib.F32 else torch.float64
    x = pos.detach().to(torch.float64).contiguous()
    v = vel.detach().to(torch.float64).contiguous()
    m = mass.detach().to(sdt).contiguous()
    r = rho.detach().to(sdt).contiguous()
    bx = None
    if box is not None:
        b = [float(t) for t in torch.as_tensor(box, dtype=torch.float64).reshape(-1).tolist()]
        bx = (C.c_double * 3)(*(b * 3 if len(b) == 1 else b))
    out = (C.c_double * ROW)()
    ctx = _lib.default_context(dev)
    _lib.check(_lib.lib().tsph_conservation_row(
        ctx.handle, _lib.ptr(x), _lib.ptr(v), _lib.ptr(m), _lib.ptr(r), x.shape[0], code,
        float(rho0), float(sound_speed), float(gamma),
        None if bx is None else C.cast(bx, C.POINTER(C.c_double)), float(wavenumber), out,
        _lib.stream_handle(dev)))
    return list(out)


def row_to_ledger(row, box=None) -> dict:
    """The dict of conservation.diagnostics from a ledger row."""
    mom = row[R_P:R_P + 3]
    if box is None:
        com = [row[R_MX + d] / row[R_MASS] for d in range(3)]
    else:
        b = [float(t) for t in torch.as_tensor(box, dtype=torch.float64).reshape(-1).tolist()]
        b = b * 3 if len(b) == 1 else b
        com = []
        for d in range(3):
            turn = math.atan2(row[R_SIN + d] / row[R_MASS], row[R_COS + d] / row[R_MASS]) / (2.0 * math.pi)
            com.append((turn - math.floor(turn)) * b[d])
    return {
        "mass": row[R_MASS],
        "momentum": list(mom),
        "momentum_magnitude": math.sqrt(sum(c * c for c in mom)),
        "angular_momentum": list(row[R_L:R_L + 3]),
        "kinetic_energy": row[R_KE],
        "internal_energy": row[R_IE],
        "total_energy": row[R_KE] + row[R_IE],
        "centre_of_mass": com,
    }


def diagnostics(pos, vel, mass, rho, *, rho0: float, sound_speed: float, gamma: float = 7.0,
                periodic_box=None, origin=None) -> dict:
    """One row of the conservation ledger (conservation.py:106-121)."""
    if origin is not None:
        pos = pos - torch.as_tensor(origin, dtype=pos.dtype, device=pos.device)
    row = conservation_row(pos, vel, mass, rho, rho0=rho0, sound_speed=sound_speed, gamma=gamma,
                           box=periodic_box)
    out = row_to_ledger(row, periodic_box)
    if origin is not None:
        # the centre of mass is reported in the caller's frame
        out["centre_of_mass"] = centre_of_mass(mass, pos + torch.as_tensor(
            origin, dtype=pos.dtype, device=pos.device), periodic_box).tolist()
    return out


def mass_total(mass: torch.Tensor) -> float:
    return float(mass.sum())


def momentum(mass: torch.Tensor, vel: torch.Tensor) -> torch.Tensor:
    return (vel * mass[:, None]).sum(0)


def angular_momentum(mass, pos, vel, origin=None) -> torch.Tensor:
    arm = pos if origin is None else pos - torch.as_tensor(origin, dtype=pos.dtype, device=pos.device)
    return (torch.linalg.cross(arm, vel) * mass[:, None]).sum(0)


def kinetic_energy(mass: torch.Tensor, vel: torch.Tensor) -> float:
    return 0.5 * float((vel.square() * mass[:, None]).sum())


def internal_energy(rho: torch.Tensor, rho0: float, sound_speed: float, gamma: float) -> torch.Tensor:
    """e(rho) = c0^2/gamma [((rho/rho0)^(gamma-1) - 1)/(gamma - 1) + rho0/rho - 1],
    the closed-form integral of -p d(1/rho) for the Tait law (conservation.py:61-75)."""
    s = rho / rho0
    return (sound_speed * sound_speed / gamma) * ((s ** (gamma - 1.0) - 1.0) / (gamma - 1.0) + 1.0 / s - 1.0)


def total_energy(mass, vel, rho, rho0: float, sound_speed: float, gamma: float) -> float:
    return kinetic_energy(mass, vel) + float((internal_energy(rho, rho0, sound_speed, gamma) * mass).sum())


def centre_of_mass(mass, pos, periodic_box=None) -> torch.Tensor:
    """Mass-weighted mean; in a periodic box the circular mean per axis
    (conservation.py:84-103)."""
    w = mass[:, None] / mass.sum()
    if periodic_box is None:
        return (pos * w).sum(0)
    L = torch.as_tensor(periodic_box, dtype=pos.dtype, device=pos.device).reshape(-1)
    L = L.expand(pos.shape[1]) if L.numel() == 1 else L
    phase = pos * (2.0 * math.pi / L)
    turn = torch.atan2((torch.sin(phase) * w).sum(0), (torch.cos(phase) * w).sum(0)) / (2.0 * math.pi)
    return turn.remainder(1.0) * L


def relative_drift(values, *, scale: float | None = None) -> dict:
    """First/last/peak summary of a scalar series (conservation.py:124-145)."""
    v = [float(x) for x in values]
    denom = abs(scale) if scale is not None else abs(v[0])
    peak = max(abs(x) for x in v)
    change = v[-1] - v[0]
    return {"initial": v[0], "final": v[-1], "delta": change, "peak_abs": peak, "scale": denom,
            "relative_drift": change / denom if denom > 0.0 else None,
            "relative_to_peak": change / peak if peak > 0.0 else None}
