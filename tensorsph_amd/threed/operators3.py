"""3-D WCSPH operators behind the reference's ``threed.operators3`` API
(operators3.py:51-429).

``accelerations`` and the single operators run on libtsph
(tsph_threed_accelerations / tsph_threed_operator): the caller's pair list is
turned into per-particle neighbour lists on the GPU, and each particle gathers
over its own list.  Precision follows the positions' dtype: float64 runs the
FP64 kernels, float32 the FP32-compensated kernels (deltas formed in FP64,
kernel arithmetic in FP32, Kahan sums); results come back in the positions'
dtype.  Tensors must live on a ROCm device.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from .. import _lib

_FORMS = {"molteni": 0, "antuono": 1}
_WALL_MODELS = {"adami": 0, "tait": 1}

OP_DENSITY, OP_PRESSURE, OP_VISCOUS, OP_ADAMI, OP_DIFFUSION = range(5)


def _code(t: torch.Tensor) -> int:
    if t.dtype == torch.float64:
        return _lib.F64
    if t.dtype == torch.float32:
        return _lib.F32
    raise TypeError(f"positions must be float32 or float64, got {t.dtype}")


def _sdt(code: int) -> torch.dtype:
    return torch.float64 if code == _lib.F64 else torch.float32


def _box3(box):
    if box is None:
        return None
    b = [float(x) for x in torch.as_tensor(box, dtype=torch.float64).reshape(-1).tolist()]
    if len(b) == 1:
        b = b * 3
    if len(b) != 3:
        raise ValueError(f"box must have 1 or 3 edges, got {len(b)}")
    return b


def _params(code, n, n_fluid, h, *, rho0=1.0, sound_speed=1.0, gamma=7.0, viscosity=0.0,
            gravity=None, box=None, renormalise=True, diffusion_scale=0.0, diffusion_form=1,
            eigenvalue_floor=0.1, dt=0.0, wall_model=0, viscous_order=2) -> _lib.ThreedParams:
    p = _lib.ThreedParams()
    p.precision = code
    p.n, p.n_fluid = int(n), int(n_fluid)
    p.h, p.rho0, p.sound_speed, p.gamma = float(h), float(rho0), float(sound_speed), float(gamma)
    p.viscosity = float(viscosity)
    if gravity is not None:
        g = [float(x) for x in torch.as_tensor(gravity, dtype=torch.float64).reshape(-1).tolist()]
        if len(g) != 3:
            raise ValueError("gravity must have 3 components")
        for d in range(3):
            p.gravity[d] = g[d]
        p.has_gravity = 1
    b = _box3(box)
    if b is not None:
        for d in range(3):
            p.box[d] = b[d]
        p.periodic = 1
    p.renormalise = 1 if renormalise else 0
    p.diffusion_scale = float(diffusion_scale)
    p.diffusion_form = int(diffusion_form)
    p.eigenvalue_floor = float(eigenvalue_floor)
    p.dt = float(dt)
    p.wall_model = int(wall_model)
    if int(viscous_order) not in (2, 4):
        raise ValueError(f"viscous_order must be 2 or 4, got {viscous_order!r}")
    p.viscous_order = int(viscous_order)
    return p


def _prepare(pos, mass, pairs, vel=None):
    """Device, precision, FP64 coordinates, scalar masses and all-target lists."""
    if pos.ndim != 2 or pos.shape[1] != 3:
        raise ValueError("pos must have shape (N, 3)")
    if pairs.ndim != 2 or pairs.shape[1] != 2:
        raise ValueError("pairs must have shape (P, 2)")
    dev = _lib.require_gpu(pos, mass, pairs, vel)
    code = _code(pos)
    n = pos.shape[0]
    if mass.shape != (n,):
        raise ValueError("mass must have shape (N,)")
    x = pos.detach().to(torch.float64).contiguous()
    v = None if vel is None else vel.detach().to(torch.float64).contiguous()
    m = mass.detach().to(_sdt(code)).contiguous()
    pr = pairs.detach().to(torch.int64).contiguous()
    ctx = _lib.default_context(dev)
    stream = _lib.stream_handle(dev)
    _lib.check(_lib.lib().tsph_nlist_from_pairs(ctx.handle, _lib.ptr(pr), pr.shape[0], n, n, 1, stream))
    return dev, code, x, v, m, ctx, stream


def _operator(op, pos, mass, h, pairs, *, vel=None, rho=None, pressure=None, out_shape, **kw):
    dev, code, x, v, m, ctx, stream = _prepare(pos, mass, pairs, vel)
    sdt = _sdt(code)
    n = pos.shape[0]
    prm = _params(code, n, kw.pop("n_fluid", n), h, **kw)
    r = None if rho is None else rho.detach().to(sdt).contiguous()
    p = None if pressure is None else pressure.detach().to(sdt).contiguous()
    out = torch.empty(out_shape, dtype=sdt, device=dev)
    _lib.check(_lib.lib().tsph_threed_operator(ctx.handle, C.byref(prm), op, _lib.ptr(x), _lib.ptr(v),
                                               _lib.ptr(m), _lib.ptr(r), _lib.ptr(p), _lib.ptr(out),
                                               stream))
    return out.to(pos.dtype)


# ----------------------------------------------------------------------- operators

def summation_density(pos: torch.Tensor, mass: torch.Tensor, h: float,
                      pairs: torch.Tensor, box=None) -> torch.Tensor:
    """rho_i = sum_j m_j W_ij + m_i W(0) (operators3.py:70-79)."""
    return _operator(OP_DENSITY, pos, mass, h, pairs, box=box, out_shape=(pos.shape[0],))


def tait_pressure(rho: torch.Tensor, rho0: float, sound_speed: float,
                  gamma: float) -> torch.Tensor:
    """max(0, B((rho/rho0)^gamma - 1)), B = rho0 c0^2/gamma, no ratio floor
    (operators3.py:82-86)."""
    dev = _lib.require_gpu(rho)
    code = _code(rho)
    r = rho.detach().contiguous()
    p = torch.empty_like(r)
    _lib.check(_lib.lib().tsph_eos_tait(_lib.ptr(r), r.numel(), code, float(rho0), float(sound_speed),
                                        float(gamma), 0.0, _lib.ptr(p), _lib.stream_handle(dev)))
    return p


def tait_density(pressure: torch.Tensor, rho0: float, sound_speed: float,
                 gamma: float) -> torch.Tensor:
    """Inverse Tait law rho0 (1 + max(p, 0)/B)^(1/gamma) (operators3.py:89-93)."""
    _lib.require_gpu(pressure)
    stiffness = rho0 * sound_speed ** 2 / gamma
    return rho0 * (1.0 + pressure.clamp_min(0.0) / stiffness).pow(1.0 / gamma)


def pressure_acceleration(pos: torch.Tensor, mass: torch.Tensor, rho: torch.Tensor,
                          pressure: torch.Tensor, h: float, pairs: torch.Tensor,
                          box=None) -> torch.Tensor:
    """-sum_j m_j (p_i/rho_i^2 + p_j/rho_j^2) grad_i W_ij (operators3.py:96-113)."""
    return _operator(OP_PRESSURE, pos, mass, h, pairs, rho=rho, pressure=pressure, box=box,
                     out_shape=(pos.shape[0], 3))


def viscous_acceleration(pos: torch.Tensor, vel: torch.Tensor, mass: torch.Tensor,
                         rho: torch.Tensor, nu: float, h: float, pairs: torch.Tensor,
                         box=None, renormalise: bool = True, order: int = 2) -> torch.Tensor:
    """nu sum_j m_j (v_j - v_i) lap W_ij / rho_ij, divided by K_ij/2 when
    renormalise (operators3.py:116-154).  order = 4 applies the iterated
    operator L v - L(beta L v), beta_ij = (I4_i + I4_j) / 240 with I4 the
    fourth moment sum_j r^4 lap W_ij m_j / rho_ij (:146-151)."""
    if order not in (2, 4):
        raise ValueError(f"order must be 2 or 4, got {order!r}")
    if nu == 0.0:
        _lib.require_gpu(pos)
        return torch.zeros_like(pos)
    return _operator(OP_VISCOUS, pos, mass, h, pairs, vel=vel, rho=rho, box=box, viscosity=nu,
                     renormalise=renormalise, viscous_order=order, out_shape=(pos.shape[0], 3))


def wall_pressure_adami(pos: torch.Tensor, mass: torch.Tensor, rho: torch.Tensor,
                        pressure: torch.Tensor, h: float, pairs: torch.Tensor, n_fluid: int,
                        gravity=None, box=None) -> torch.Tensor:
    """Adami extrapolated pressure of the wall rows n_fluid.. (operators3.py:173-209)."""
    n = pos.shape[0]
    if n_fluid >= n:
        _lib.require_gpu(pos)
        return torch.zeros(0, dtype=pos.dtype, device=pos.device)
    return _operator(OP_ADAMI, pos, mass, h, pairs, rho=rho, pressure=pressure, box=box,
                     gravity=gravity, n_fluid=n_fluid, out_shape=(n - n_fluid,))


def density_diffusion_rate_3d(pos: torch.Tensor, rho: torch.Tensor, mass: torch.Tensor, h: float,
                              pairs: torch.Tensor, *, sound_speed: float, delta: float,
                              form: str = "antuono", eigenvalue_floor: float = 0.1,
                              n_fluid: int | None = None, box=None) -> torch.Tensor:
    """Fluid-fluid diffusive density rate, scale 2 delta c0 h (operators3.py:259-295)."""
    if delta <= 0.0:
        _lib.require_gpu(rho)
        return torch.zeros_like(rho)
    if form not in _FORMS:
        raise ValueError(f"unknown density_diffusion_form {form!r}; expected one of "
                         f"{tuple(_FORMS)}")
    n = pos.shape[0]
    return _operator(OP_DIFFUSION, pos, mass, h, pairs, rho=rho, box=box,
                     n_fluid=n if n_fluid is None else n_fluid, sound_speed=sound_speed,
                     diffusion_scale=2.0 * delta * sound_speed * h, diffusion_form=_FORMS[form],
                     eigenvalue_floor=eigenvalue_floor, dt=1.0, out_shape=(n,))


def accelerations(pos: torch.Tensor, vel: torch.Tensor, mass: torch.Tensor, h: float,
                  pairs: torch.Tensor, *, rho0: float, sound_speed: float, gamma: float = 7.0,
                  viscosity: float = 0.0, gravity=None, box=None, n_fluid: int | None = None,
                  renormalise: bool = True, wall_model: str = "adami", viscous_order: int = 2,
                  density_diffusion: float = 0.0, dt: float | None = None,
                  density_diffusion_form: str = "antuono",
                  density_diffusion_eigenvalue_floor: float = 0.1):
    """Full right-hand side (operators3.py:212-256): returns (acc, rho, pressure)."""
    n = pos.shape[0]
    n_fluid = n if n_fluid is None else int(n_fluid)
    if density_diffusion > 0.0 and (dt is None or dt <= 0.0):
        raise ValueError("density_diffusion > 0 requires a positive dt")
    if n_fluid < n and wall_model not in _WALL_MODELS:
        raise ValueError(f"unknown wall_model {wall_model!r}")
    if viscous_order not in (2, 4):
        raise ValueError(f"viscous_order must be 2 or 4, got {viscous_order!r}")
    if density_diffusion > 0.0 and density_diffusion_form not in _FORMS:
        raise ValueError(f"unknown density_diffusion_form {density_diffusion_form!r}; "
                         f"expected one of {tuple(_FORMS)}")
    dev, code, x, v, m, ctx, stream = _prepare(pos, mass, pairs, vel)
    sdt = _sdt(code)
    prm = _params(code, n, n_fluid, h, rho0=rho0, sound_speed=sound_speed, gamma=gamma,
                  viscosity=viscosity, gravity=gravity, box=box, renormalise=renormalise,
                  diffusion_scale=(2.0 * density_diffusion * sound_speed * h
                                   if density_diffusion > 0.0 else 0.0),
                  diffusion_form=_FORMS.get(density_diffusion_form, 1),
                  eigenvalue_floor=density_diffusion_eigenvalue_floor,
                  dt=0.0 if dt is None else dt,
                  wall_model=_WALL_MODELS.get(wall_model, 0), viscous_order=viscous_order)
    acc = torch.empty((n, 3), dtype=sdt, device=dev)
    rho = torch.empty(n, dtype=sdt, device=dev)
    p = torch.empty(n, dtype=sdt, device=dev)
    _lib.check(_lib.lib().tsph_threed_accelerations(ctx.handle, C.byref(prm), _lib.ptr(x), _lib.ptr(v),
                                                    _lib.ptr(m), _lib.ptr(acc), _lib.ptr(rho),
                                                    _lib.ptr(p), stream))
    return acc.to(pos.dtype), rho.to(pos.dtype), p.to(pos.dtype)


# ----------------------------------------------------------------------- shear-wave helpers

def mode_amplitude(pos: torch.Tensor, vel: torch.Tensor, wavenumber: float) -> float:
    """4 <v_x sin(k y) sin(k z)> (operators3.py:401-408)."""
    shape = torch.sin(wavenumber * pos[:, 1]) * torch.sin(wavenumber * pos[:, 2])
    return float(4.0 * (vel[:, 0] * shape).mean())


def analytic_mode_decay(nu: float, wavenumber: float) -> float:
    """2 nu k^2, the decay exponent of u = A sin(k y) sin(k z) (operators3.py:411-413)."""
    return 2.0 * nu * wavenumber ** 2


def shear_wave_field(pos: torch.Tensor, amplitude: float, wavenumber: float) -> torch.Tensor:
    """u_x = A sin(k y) sin(k z) (operators3.py:416-421)."""
    vel = torch.zeros_like(pos)
    vel[:, 0] = amplitude * torch.sin(wavenumber * pos[:, 1]) * torch.sin(wavenumber * pos[:, 2])
    return vel


def cell_dx(box: float, n: int) -> float:
    """box / n (operators3.py:424-426)."""
    return float(box) / float(n)


def wavenumber_for(box: float, periods: int = 1) -> float:
    """2 pi periods / box (operators3.py:429-431)."""
    return 2.0 * math.pi * periods / float(box)
