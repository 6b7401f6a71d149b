"""The C-ABI library without a GPU: it loads, exports every entry point that
include/pycles_amd.h declares, and rejects bad arguments before any device
call (argument checks run on the host, so these calls never touch a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pycles_amd.h")
LIB = os.path.join(ROOT, "pycles_amd", "libpycles_amd.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libpycles_amd.so not built")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void|char)\s*\*?\s*(pycles_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    return C.CDLL(LIB)


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("pycles_advection", "pycles_thermo_sa", "pycles_thermo_dry", "pycles_rk3_stage",
              "pycles_pressure_plan_create", "pycles_pressure_update", "pycles_pressure_plan_destroy"):
        assert n in names


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


class Dims(C.Structure):
    _fields_ = [("n", C.c_int64 * 3), ("gw", C.c_int64), ("dx", C.c_double * 3), ("dxi", C.c_double * 3)]


class Ref(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("rho0", "rho0_half", "alpha0", "alpha0_half", "p0", "p0_half")]


def _dims(n=(16, 16, 8), gw=3):
    d = Dims()
    for i in range(3):
        d.n[i] = n[i]
        d.dx[i] = 10.0
        d.dxi[i] = 0.1
    d.gw = gw
    return d


def test_version_and_error_text(lib):
    lib.pycles_version.restype = C.c_char_p
    lib.pycles_last_error.restype = C.c_char_p
    assert lib.pycles_version().decode()
    assert isinstance(lib.pycles_last_error(), bytes)


def _advection(lib, dims, so, mo, nscal=1, precision=0):
    fake = C.c_void_p(0x1000)      # never dereferenced: the checks fail first
    ref = Ref(*([fake.value] * 6))
    vel = (C.c_void_p * 3)(fake.value, fake.value, fake.value)
    tv = (C.c_void_p * 3)(fake.value, fake.value, fake.value)
    sc = (C.c_void_p * 2)(fake.value, fake.value)
    ts = (C.c_void_p * 2)(fake.value, fake.value)
    lib.pycles_advection.restype = C.c_int
    return lib.pycles_advection(C.byref(dims), C.byref(ref), vel, tv, nscal, sc, ts, None, so, mo,
                                precision, 0, None)


def test_advection_rejects_bad_arguments(lib):
    lib.pycles_last_error.restype = C.c_char_p
    assert _advection(lib, _dims(), 6, 5) == -3           # order 6: unsupported
    assert b"order" in lib.pycles_last_error()
    assert _advection(lib, _dims(gw=3), 9, 5) == -1       # WENO9 needs gw >= 5
    assert _advection(lib, _dims(n=(5, 16, 8)), 5, 5) == -1   # nx < 2r
    assert _advection(lib, _dims(), 5, 5, nscal=3) == -1
    assert _advection(lib, _dims(n=(16, 16, 0)), 5, 5) == -1
    assert _advection(lib, _dims(), 5, 5, precision=7) == -1


def test_pressure_plan_rejects_unsupported_sizes(lib):
    lib.pycles_pressure_plan_create.restype = C.c_int
    rho = (C.c_double * 64)(*([1.0] * 64))
    plan = C.c_void_p()
    rc = lib.pycles_pressure_plan_create(C.byref(_dims(n=(12, 16, 8))), rho, rho, 0, C.byref(plan))
    assert rc != 0 and not plan.value
    # the z solve holds its level profiles in LDS: at most 2048 levels
    lib.pycles_last_error.restype = C.c_char_p
    rc = lib.pycles_pressure_plan_create(C.byref(_dims(n=(16, 16, 2049))), rho, rho, 0, C.byref(plan))
    assert rc == -3 and not plan.value
    assert b"2048" in lib.pycles_last_error()


def _chunks(lib, kernel, n, gw=3):
    nch, ln = C.c_int(), C.c_int()
    lib.pycles_launch_chunks.restype = C.c_int
    rc = lib.pycles_launch_chunks(kernel, C.byref(_dims(n=n, gw=gw)), C.byref(nch), C.byref(ln))
    assert rc == 0
    return nch.value, ln.value


def test_z_chunks_partition_every_level_count(lib):
    """The z chunks of the chunked kernels, as their grids use them (the
    formulae in the header), cover each level exactly once with no empty
    chunk; the pressure correction's chunks hold at least gw levels each
    (its first and last chunks fill the ghost levels from them); the fused
    y-z pass's chunks hold at most the 128 levels its LDS profile copy
    holds."""
    for gw in (3, 5):
        for plane in ((64, 64), (512, 512), (16, 8)):
            for nz in range(max(2, gw), 2049):
                n = (plane[0], plane[1], nz)
                nch, ln = _chunks(lib, 0, n, gw)
                ranges = [(gw + b * ln, nz + gw if b + 1 == nch else gw + (b + 1) * ln) for b in range(nch)]
                assert ranges[0][0] == gw and ranges[-1][1] == nz + gw
                assert all(r[1] == q[0] for r, q in zip(ranges, ranges[1:]))
                lens = [r[1] - r[0] for r in ranges]
                assert min(lens) >= (gw if nch > 1 else 1), (gw, n, ranges)
                for kern, total in ((1, plane[1]), (2, nz), (3, nz + 2 * gw), (4, nz)):
                    nch, ln = _chunks(lib, kern, n, gw)
                    ranges = [(b * ln, min((b + 1) * ln, total)) for b in range(nch)]
                    assert ranges[-1][1] == total and all(r[1] > r[0] for r in ranges), (kern, n, ranges)
                    assert kern != 4 or ln <= 128, (n, ln)


def test_layout_and_cfl_read_reject_bad_arguments(lib):
    lib.pycles_layout_ingress.restype = C.c_int
    lib.pycles_layout_egress.restype = C.c_int
    lib.pycles_cfl_read.restype = C.c_int
    fake = C.c_void_p(0x1000)      # never dereferenced: the checks fail first
    d = _dims()
    assert lib.pycles_layout_ingress(C.byref(d), None, fake, None, 0, None) == -1
    assert lib.pycles_layout_ingress(C.byref(d), fake, fake, None, 3, None) == -1
    assert lib.pycles_layout_egress(C.byref(d), fake, None, None, 0, None) == -1
    assert lib.pycles_layout_egress(C.byref(d), fake, fake, None, -1, None) == -1
    assert lib.pycles_cfl_read(None, None, None) == -1
    assert lib.pycles_launch_chunks(9, C.byref(d), C.byref(C.c_int()), C.byref(C.c_int())) == -1
