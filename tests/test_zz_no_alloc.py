"""The allocation-free conservation row (tsph_conservation_row_noalloc)
against tsph_conservation_row, through the C ABI: the same 21 doubles bit for
bit on two calls in a row, the second with fewer rows, in both precisions;
sizes beyond its fixed device array are refused before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ROW = 21
MAX_ROWS = 12_782_336


def _bind():
    from tensorsph_amd import _lib
    L = _lib.lib()
    f = L.tsph_conservation_row_noalloc
    f.argtypes = [C.c_void_p] * 4 + [C.c_int64, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double),
                                     C.c_void_p]
    f.restype = C.c_int
    return _lib, L


def _state(n, dtype, seed):
    g = torch.Generator(device="cuda:0").manual_seed(seed)
    dev = torch.device("cuda:0")
    pos = torch.rand((n, 3), generator=g, device=dev, dtype=torch.float64)
    vel = torch.randn((n, 3), generator=g, device=dev, dtype=torch.float64) * 1e-3
    mass = (torch.rand(n, generator=g, device=dev, dtype=torch.float64) + 0.5).to(dtype)
    rho = (1.0 + 0.01 * torch.randn(n, generator=g, device=dev, dtype=torch.float64)).to(dtype)
    return pos, vel, mass, rho


def _rows(n, dtype, seed):
    _lib, L = _bind()
    pos, vel, mass, rho = _state(n, dtype, seed)
    code = _lib.F64 if dtype == torch.float64 else _lib.F32
    box = (C.c_double * 3)(1.0, 1.0, 1.0)
    stream = _lib.stream_handle(pos.device)
    ref = (C.c_double * ROW)()
    ctx = _lib.default_context(pos.device)
    _lib.check(L.tsph_conservation_row(ctx.handle, _lib.ptr(pos), _lib.ptr(vel), _lib.ptr(mass),
                                       _lib.ptr(rho), n, code, 1.0, 10.0, 7.0, box, 2 * np.pi, ref,
                                       stream))
    got = (C.c_double * ROW)()
    _lib.check(L.tsph_conservation_row_noalloc(_lib.ptr(pos), _lib.ptr(vel), _lib.ptr(mass),
                                               _lib.ptr(rho), n, code, 1.0, 10.0, 7.0, box,
                                               2 * np.pi, got, stream))
    return np.array(ref[:]), np.array(got[:])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_row_equals_the_allocating_one(dtype):
    for n, seed in ((300_001, 1), (4_099, 2)):          # the second call is smaller
        ref, got = _rows(n, dtype, seed)
        assert np.array_equal(ref, got), (n, np.abs(ref - got).max())
        assert ref[0] > 0.0


def test_noalloc_row_refuses_more_rows_than_its_array():
    _lib, L = _bind()
    t = torch.zeros((4, 3), dtype=torch.float64, device="cuda:0")
    out = (C.c_double * ROW)()
    st = L.tsph_conservation_row_noalloc(_lib.ptr(t), _lib.ptr(t), _lib.ptr(t), _lib.ptr(t),
                                         MAX_ROWS + 1, _lib.F64, 1.0, 10.0, 7.0, None, 0.0, out,
                                         _lib.stream_handle(t.device))
    assert st == _lib.TSPH_ERR_ARG
    assert b"exceed" in L.tsph_last_error()
