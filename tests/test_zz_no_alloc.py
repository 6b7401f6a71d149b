"""pycles_horizontal_stats_noalloc, the statistics entry point that allocates
nothing, against pycles_horizontal_stats (workspace argument) through the C
ABI: bitwise equal per-level moments, extremes and counts, in FP64 and FP32,
plain, as a product with level profiles, and masked, on two calls in a row
whose second is smaller (the library's array keeps the first call's
partials).  Sizes beyond the array are refused on the host."""
import ctypes as C
import os

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pycles_amd", "libpycles_amd.so")

pytestmark = pytest.mark.gpu

F64, F32 = 0, 1


class Dims(C.Structure):
    _fields_ = [("n", C.c_int64 * 3), ("gw", C.c_int64), ("dx", C.c_double * 3), ("dxi", C.c_double * 3)]


def _dims(nx, ny, nz, gw):
    d = Dims()
    for i, n in enumerate((nx, ny, nz)):
        d.n[i] = n
        d.dx[i] = 20.0
        d.dxi[i] = 0.05
    d.gw = gw
    return d


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libpycles_amd.so is not built")
    L = C.CDLL(LIB)
    vp = C.c_void_p
    L.pycles_horizontal_stats.argtypes = [C.POINTER(Dims), vp, vp, vp, vp, vp, C.c_int, vp, vp, vp]
    L.pycles_horizontal_stats_noalloc.argtypes = [C.POINTER(Dims), vp, vp, vp, vp, vp, C.c_int, vp, vp]
    L.pycles_horizontal_stats_workspace_bytes.argtypes = [C.POINTER(Dims)]
    L.pycles_horizontal_stats_workspace_bytes.restype = C.c_size_t
    L.pycles_last_error.restype = C.c_char_p
    return L


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _case(nx, ny, nz, gw, dtype, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    nzg = nz + 2 * gw
    a = (torch.randn(nzg, ny, nx, generator=g, dtype=torch.float64) * 3.0).to(dtype).cuda()
    b = (torch.randn(nzg, ny, nx, generator=g, dtype=torch.float64) + 2.0).to(dtype).cuda()
    mask = (torch.rand(nzg, ny, nx, generator=g, dtype=torch.float64) > 0.6).to(dtype).cuda()
    base_a = torch.linspace(300.0, 310.0, nzg, dtype=torch.float64).cuda()
    base_b = torch.linspace(0.01, 0.02, nzg, dtype=torch.float64).cuda()
    return a, b, mask, base_a, base_b


def _both(lib, d, args, precision):
    nzg = int(d.n[2] + 2 * d.gw)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    out_ws = torch.full((6 * nzg,), np.nan, dtype=torch.float64, device="cuda")
    out_na = torch.full((6 * nzg,), np.nan, dtype=torch.float64, device="cuda")
    ws = torch.empty(lib.pycles_horizontal_stats_workspace_bytes(C.byref(d)), dtype=torch.uint8, device="cuda")
    a, base_a, b, base_b, mask = args
    rc = lib.pycles_horizontal_stats(C.byref(d), _ptr(a), _ptr(base_a), _ptr(b), _ptr(base_b), _ptr(mask),
                                     precision, _ptr(out_ws), _ptr(ws), stream)
    assert rc == 0, lib.pycles_last_error()
    rc = lib.pycles_horizontal_stats_noalloc(C.byref(d), _ptr(a), _ptr(base_a), _ptr(b), _ptr(base_b), _ptr(mask),
                                             precision, _ptr(out_na), stream)
    assert rc == 0, lib.pycles_last_error()
    torch.cuda.synchronize()
    return out_ws.cpu().numpy(), out_na.cpu().numpy()


@pytest.mark.parametrize("precision", [F64, F32])
def test_noalloc_matches_workspace_entry_point(gpu, lib, precision):
    dtype = torch.float64 if precision == F64 else torch.float32
    # a larger call, then a smaller one (fewer levels and columns)
    for (nx, ny, nz, gw), seed in (((48, 40, 120, 3), 5), ((16, 8, 10, 5), 6)):
        print(f"noalloc stats: precision {precision}, {nx}x{ny}x{nz} gw {gw}", flush=True)
        d = _dims(nx, ny, nz, gw)
        a, b, mask, base_a, base_b = _case(nx, ny, nz, gw, dtype, seed)
        base = (base_a, base_b) if precision == F32 else (None, None)
        for args in ((a, base[0], None, None, None),
                     (a, base[0], b, base[1], None),
                     (a, base[0], None, None, mask)):
            ref, got = _both(lib, d, args, precision)
            assert np.all(np.isfinite(ref)), "workspace entry point left unwritten outputs"
            assert np.array_equal(ref, got), np.abs(ref - got).max()


def test_noalloc_refuses_levels_beyond_its_array(lib):
    # host-side check: returns before any device work
    d = _dims(4, 4, 2043, 3)
    rc = lib.pycles_horizontal_stats_noalloc(C.byref(d), C.c_void_p(16), None, None, None, None, F64,
                                             C.c_void_p(16), None)
    assert rc == -3 and b"2048" in lib.pycles_last_error()
