"""ctypes binding of libtsph.so (the C ABI declared in include/tsph.h).

The library is built in-tree by ``__graft_entry__.build()`` /
``make -C tensorsph_amd/csrc``.  There is no CPU fallback: every public
operator of tensorsph_amd runs on the GPU and fails loudly when the library
or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libtsph.so"

TSPH_OK, TSPH_ERR_ARG, TSPH_ERR_HIP, TSPH_ERR_OOM, TSPH_ERR_NUMERIC, TSPH_ERR_TYPE = range(6)
F64, F32 = 0, 1

_lib = None
_lock = threading.Lock()


class LibraryMissing(RuntimeError):
    pass


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not LIB_PATH.exists():
                    raise LibraryMissing(
                        f"{LIB_PATH} is not built; run `make -C tensorsph_amd/csrc` "
                        "(or __graft_entry__.build()) first")
                L = C.CDLL(str(LIB_PATH))
                _declare(L)
                _lib = L
    return _lib


_vp = C.c_void_p
_i64 = C.c_int64
_dbl = C.c_double


class WcsphParams(C.Structure):
    _fields_ = [("dim", C.c_int32), ("precision", C.c_int32), ("n_fluid", C.c_int64),
                ("n_wall", C.c_int64), ("h", _dbl), ("rho0", _dbl), ("sound_speed", _dbl),
                ("gamma", _dbl), ("gravity", _dbl), ("viscosity", _dbl),
                ("wall_friction", _dbl), ("dt", _dbl), ("diffusion_scale", _dbl),
                ("diffusion_form", C.c_int32), ("eigenvalue_floor", _dbl),
                ("reaction_style", C.c_int32), ("has_wall_accel", C.c_int32),
                ("wrap_width", _dbl), ("density_renormalization", C.c_int32),
                ("lattice_kernel_sum", _dbl), ("boundary_consistent_density", C.c_int32),
                ("surface_pressure_closure", C.c_int32), ("surface_pressure_threshold", _dbl),
                ("surface_pressure_donors_all", C.c_int32), ("surface_pressure_layers", C.c_int32)]


class ThreedParams(C.Structure):
    _fields_ = [("precision", C.c_int32), ("n", C.c_int64), ("n_fluid", C.c_int64),
                ("h", _dbl), ("rho0", _dbl), ("sound_speed", _dbl), ("gamma", _dbl),
                ("viscosity", _dbl), ("gravity", _dbl * 3), ("has_gravity", C.c_int32),
                ("box", _dbl * 3), ("periodic", C.c_int32), ("renormalise", C.c_int32),
                ("diffusion_scale", _dbl), ("diffusion_form", C.c_int32),
                ("eigenvalue_floor", _dbl), ("dt", _dbl), ("wall_model", C.c_int32),
                ("viscous_order", C.c_int32)]


class GaugeSpec(C.Structure):
    _fields_ = [("station_x", _dbl), ("aperture", _dbl), ("tank_depth", _dbl), ("dx", _dbl),
                ("h", _dbl), ("gravity", _dbl), ("obstacle_lo", _dbl * 3), ("obstacle_hi", _dbl * 3),
                ("per_tap", C.c_int32), ("nodes", _dbl * (8 * 7 * 3))]


GAUGE_ROW = 117


class TvcParams(C.Structure):
    _fields_ = [("alpha", _dbl), ("limiter_slope", _dbl), ("max_shift", _dbl), ("wall_volume", _dbl),
                ("scope", C.c_int32), ("bulk_c", _dbl), ("bulk_g", _dbl), ("percentile", _dbl)]


class ShiftParams(C.Structure):
    _fields_ = [(k, _dbl) for k in ("dt", "v_ref", "v_floor", "strength", "limit", "free_surface",
                                    "percentile", "c_reference")]


class DeltaPlusParams(C.Structure):
    _fields_ = [("precision", C.c_int32), ("n", C.c_int64), ("n_wall", C.c_int64),
                ("dx", _dbl), ("sound_speed", _dbl), ("reference_speed", _dbl), ("dt", _dbl),
                ("rho0", _dbl), ("gravity", _dbl), ("viscosity", _dbl), ("diffusion", _dbl),
                ("alpha", _dbl)]


class BatchCaps(C.Structure):
    _fields_ = [("fluid", _i64), ("wall", _i64), ("sims", _i64), ("cells", _i64),
                ("fluid_entries", _i64), ("wall_entries", _i64), ("sorted_lists", C.c_int32)]


class DpCase(C.Structure):
    _fields_ = [("n", _i64), ("n_wall", _i64), ("steps", _i64), ("dx", _dbl), ("sound_speed", _dbl),
                ("reference_speed", _dbl), ("rho0", _dbl), ("gravity", _dbl), ("viscosity", _dbl),
                ("diffusion", _dbl), ("alpha", _dbl), ("courant", _dbl)]


class SimParams(C.Structure):
    _fields_ = [("op", WcsphParams), ("skin", _dbl), ("lo", _dbl * 3), ("hi", _dbl * 3),
                ("tol", _dbl), ("clamp", C.c_int32), ("obstacle_begin", _i64), ("obstacle_end", _i64),
                ("obstacle_centre", _dbl * 3), ("obstacle_clamp", C.c_int32),
                ("obstacle_lo", _dbl * 3), ("obstacle_hi", _dbl * 3), ("obstacle_eps", _dbl),
                ("adaptive", C.c_int32), ("target_time", _dbl), ("courant_num", _dbl),
                ("dt_limit", _dbl), ("dt_safety", _dbl), ("dt_growth", _dbl),
                ("adaptive_free_last", C.c_int32), ("adaptive_tol", _dbl),
                ("initial_time", _dbl), ("initial_dt", _dbl)]


_dp = C.POINTER(_dbl)
_ip = C.POINTER(_i64)


def _declare(L):
    L.tsph_last_error.restype = C.c_char_p
    L.tsph_abi_version.restype = C.c_int
    sigs = {
        "tsph_ctx_create": [C.POINTER(_vp)],
        "tsph_ctx_destroy": [_vp],
        "tsph_pairs_search": [_vp, _vp, C.c_int, _i64, C.c_int, _dbl, C.POINTER(_dbl), _i64, _vp,
                              C.POINTER(_i64)],
        "tsph_pairs_brute": [_vp, _vp, C.c_int, _i64, C.c_int, _dbl, C.POINTER(_dbl), _vp,
                             C.POINTER(_i64)],
        "tsph_pairs_export": [_vp, _vp, _vp],
        "tsph_cells_export": [_vp, _vp, _vp, _vp, C.POINTER(_i64), _vp],
        "tsph_max_displacement": [_vp, _vp, C.c_int, _i64, C.c_int, _vp, _vp],
        "tsph_nlist_from_pairs": [_vp, _vp, _i64, _i64, _i64, C.c_int, _vp],
        "tsph_wcsph_rhs": [_vp, C.POINTER(WcsphParams), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                           _vp, _vp, _vp],
        "tsph_wcsph_rhs_ex": [_vp, C.POINTER(WcsphParams), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                              _vp, _vp, _vp, _vp, _vp],
        "tsph_gauges": [_vp, C.POINTER(GaugeSpec), C.c_int, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
        "tsph_sim_gauges": [_vp, C.POINTER(GaugeSpec), _vp, _vp],
        "tsph_sim_shift": [_vp, C.POINTER(ShiftParams), _vp, _vp],
        "tsph_sim_tvc": [_vp, C.POINTER(TvcParams), _vp, _vp],
        "tsph_eos_tait": [_vp, _i64, C.c_int, _dbl, _dbl, _dbl, _dbl, _vp, _vp],
        "tsph_kernel_eval": [C.c_int, _vp, _i64, C.c_int, _dbl, _vp, _vp, _vp, _vp],
        "tsph_apply_boundary": [_vp, _vp, _i64, C.c_int, _dp, _dp, _dbl, _vp, _vp],
        "tsph_obstacle_clamp": [_vp, _vp, _vp, C.c_int, _i64, _dp, _dp, _dbl, _vp, _vp, _vp],
        "tsph_sim_create": [C.POINTER(SimParams), _vp, _vp, _vp, _vp, C.POINTER(_vp)],
        "tsph_sim_step": [_vp, _i64, _i64, _vp],
        "tsph_sim_read": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
        "tsph_sim_stats": [_vp, _ip, _dp],
        "tsph_sim_meters": [_vp, _dp, _i64, _ip],
        "tsph_sim_loads": [_vp, _dp, _i64, _ip],
        "tsph_sim_time": [_vp, _dp, _vp],
        "tsph_sim_timing": [_vp, C.c_int, _dp, _ip],
        "tsph_sim_configure": [_vp, C.c_int, C.c_int, C.c_int],
        "tsph_sim_set_masses_f64": [_vp, _vp, _vp],
        "tsph_sim_layout": [_vp, _ip],
        "tsph_sim_slot_width": [_vp, _i64, _ip],
        "tsph_sim_count_pairs": [_vp, _dbl, _ip, _vp],
        "tsph_sim_destroy": [_vp],
        "tsph_threed_accelerations": [_vp, C.POINTER(ThreedParams), _vp, _vp, _vp, _vp, _vp, _vp, _vp],
        "tsph_threed_operator": [_vp, C.POINTER(ThreedParams), C.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                 _vp],
        "tsph_psim_create": [C.POINTER(ThreedParams), _dbl, _vp, _vp, _vp, _vp, C.POINTER(_vp)],
        "tsph_psim_step": [_vp, _i64, _vp],
        "tsph_psim_set_masses_f64": [_vp, _vp, _vp],
        "tsph_psim_diagnostics": [_vp, _dbl, _dp, _vp],
        "tsph_psim_read": [_vp, _vp, _vp, _vp, _vp, _vp],
        "tsph_psim_stats": [_vp, _ip],
        "tsph_psim_destroy": [_vp],
        "tsph_deltaplus_step": [_vp, C.POINTER(DeltaPlusParams)] + [_vp] * 14,
        "tsph_batch_create": [C.POINTER(BatchCaps), _vp, C.POINTER(_vp)],
        "tsph_batch_admit": [_vp, C.POINTER(DpCase)] + [_vp] * 7 + [C.POINTER(C.c_int32), _vp],
        "tsph_batch_step": [_vp, _i64, _vp],
        "tsph_batch_launch": [_vp, C.c_int32, C.c_int32, C.c_int32],
        "tsph_batch_poll": [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32,
                            C.POINTER(C.c_int32)],
        "tsph_batch_read": [_vp, C.c_int32, _vp, _vp, _vp, C.POINTER(_i64), _vp],
        "tsph_batch_release": [_vp, C.c_int32, _vp],
        "tsph_batch_stats": [_vp, _ip],
        "tsph_batch_destroy": [_vp],
        "tsph_conservation_row": [_vp, _vp, _vp, _vp, _vp, _i64, C.c_int, _dbl, _dbl, _dbl, _dp, _dbl, _dp,
                                  _vp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    for name, args, res in _EXTRA:
        if hasattr(L, name):
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res


# entry points added by later modules register themselves here
_EXTRA: list = []


def register(name, args, res=C.c_int):
    _EXTRA.append((name, args, res))
    if _lib is not None and hasattr(_lib, name):
        f = getattr(_lib, name)
        f.argtypes = args
        f.restype = res


def check(status: int) -> None:
    if status == TSPH_OK:
        return
    msg = lib().tsph_last_error().decode(errors="replace")
    if status == TSPH_ERR_ARG:
        raise ValueError(msg)
    if status == TSPH_ERR_TYPE:
        raise TypeError(msg)
    if status == TSPH_ERR_NUMERIC:
        raise FloatingPointError(msg)
    if status == TSPH_ERR_OOM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


class Context:
    """Owns one tsph_ctx (device scratch: cell grid, neighbour lists)."""

    def __init__(self):
        h = _vp()
        check(lib().tsph_ctx_create(C.byref(h)))
        self.handle = h

    def __del__(self):
        try:
            if self.handle and _lib is not None:
                _lib.tsph_ctx_destroy(self.handle)
        except Exception:
            pass
        self.handle = None


_ctx_local = threading.local()


def default_context(device: torch.device) -> Context:
    key = device.index if device.index is not None else torch.cuda.current_device()
    store = getattr(_ctx_local, "store", None)
    if store is None:
        store = _ctx_local.store = {}
    if key not in store:
        with torch.cuda.device(key):
            store[key] = Context()
    return store[key]


def stream_handle(device: torch.device) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t: torch.Tensor | None) -> C.c_void_p:
    return C.c_void_p(0) if t is None else C.c_void_p(t.data_ptr())


def require_gpu(*tensors) -> torch.device:
    dev = None
    for t in tensors:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor):
            continue
        if t.device.type != "cuda":
            raise ValueError(
                "tensorsph_amd runs on the GPU: move the tensors to a 'cuda' (ROCm) device "
                f"(got {t.device})")
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise ValueError("all tensors must be on the same device")
    if dev is None:
        raise ValueError("no tensor argument to take the device from")
    return dev
