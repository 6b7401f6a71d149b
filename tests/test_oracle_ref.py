"""The oracle against the reference's own C, compiled from the reference tree
into oracle/_ref (``make -C oracle ref``): bit-for-bit agreement of the WENO
and centred interpolations, the whole momentum advection, the scalar flux
divergence, and the boundary fill (periodic x/y exchange on one rank plus
set_bcs in z).  Skipped where the reference tree is absent.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import reflib as R

pytestmark = pytest.mark.skipif(not R.available(), reason="reference tree not present")


def _stencils(rng, n, width):
    """Random stencils of mixed character: smooth, noisy, steps, tiny and huge scales."""
    out = []
    x = np.arange(width) - width // 2
    for m in range(n):
        kind = m % 5
        if kind == 0:
            q = rng.normal(size=width)
        elif kind == 1:
            q = np.sin(0.3 * x + rng.uniform(0, 6)) * rng.uniform(0.1, 10)
        elif kind == 2:
            q = np.where(x < rng.integers(-width // 2, width // 2 + 1), 1.0, -0.5) * rng.uniform(1, 300)
        elif kind == 3:
            q = rng.normal(size=width) * 1e-6 + rng.uniform(280, 320)
        else:
            q = rng.normal(size=width) * 10.0 ** rng.uniform(-8, 4)
        out.append(q)
    return out


@pytest.mark.parametrize("order", [5, 7, 9])
def test_weno_bitwise(order):
    rng = np.random.default_rng(order)
    name = f"interp_weno{order}"
    for q in _stencils(rng, 4000, order):
        a = O.weno(order, q)
        b = R.interp(name, *q)
        assert a == b or (np.isnan(a) and np.isnan(b)), (q, a, b)


@pytest.mark.parametrize("order,name", [(5, "interp_4"), (7, "interp_6"), (9, "interp_8")])
def test_centred_bitwise(order, name):
    rng = np.random.default_rng(order + 100)
    width = order - 1
    for q in _stencils(rng, 2000, width):
        assert O.centred(order, q) == R.interp(name, *q)


def _profiles(rng, nlg2):
    rho0 = rng.uniform(0.6, 1.2, nlg2)
    rho0_half = rng.uniform(0.6, 1.2, nlg2)
    return rho0, rho0_half, 1.0 / rho0, 1.0 / rho0_half


@pytest.mark.parametrize("order", [5, 7, 9])
def test_momentum_advection_bitwise(order):
    gw = (order + 1) // 2
    nx, ny, nz = 9, 7, 8
    dx = (40.0, 55.0, 17.5)
    rng = np.random.default_rng(10 + order)
    od = O.make_dims(nx, ny, nz, gw, *dx)
    rd = R.make_dims(nx, ny, nz, gw, *dx)
    shape = tuple(od.nlg)
    vel = [np.ascontiguousarray(rng.normal(size=shape) * s + o)
           for s, o in ((3.0, 1.0), (2.0, -0.5), (1.0, 0.0))]
    rho0, rho0_half, alpha0, alpha0_half = _profiles(rng, shape[2])
    want = R.momentum_tendencies(rd, rho0, rho0_half, alpha0, alpha0_half, vel, order)

    ref = O.Ref(np.ones(shape[2]), np.ones(shape[2]), alpha0, alpha0_half, rho0, rho0_half)
    sim = O.OracleSimulation(od, ref, False, None, momentum_order=order)
    for n, v in zip(("u", "v", "w"), vel):
        sim.pv(n)[...] = v
    sim.momentum_advection_update()
    for a, n in enumerate(("u", "v", "w")):
        got = sim.pt(n)
        assert np.array_equal(got, want[a]), f"{n}: max diff {np.abs(got - want[a]).max()}"


@pytest.mark.parametrize("d", [0, 1, 2])
def test_scalar_flux_divergence_bitwise(d):
    import ctypes as C
    nx, ny, nz, gw = 8, 6, 10, 3
    rng = np.random.default_rng(20 + d)
    od = O.make_dims(nx, ny, nz, gw, 30.0, 45.0, 12.0)
    rd = R.make_dims(nx, ny, nz, gw, 30.0, 45.0, 12.0)
    shape = tuple(od.nlg)
    flux = rng.normal(size=shape)
    base = rng.normal(size=shape)
    _, _, alpha0, alpha0_half = _profiles(rng, shape[2])
    t_ref = base.copy()
    R.scalar_flux_divergence(rd, alpha0, alpha0_half, flux, t_ref, d)
    t_orc = base.copy()
    O.lib().orc_scalar_flux_divergence(C.byref(od), O.ptr(alpha0), O.ptr(alpha0_half), O.ptr(flux),
                                       O.ptr(t_orc), d)
    assert np.array_equal(t_orc, t_ref)


@pytest.mark.parametrize("bc", [1.0, -1.0, 2.0])
@pytest.mark.parametrize("gw", [3, 5])
def test_boundary_fill_bitwise(bc, gw):
    import ctypes as C
    nx, ny, nz = 7, 9, 6
    rng = np.random.default_rng(int(30 + 10 * bc + gw))
    od = O.make_dims(nx, ny, nz, gw, 1.0, 1.0, 1.0)
    rd = R.make_dims(nx, ny, nz, gw, 1.0, 1.0, 1.0)
    v = rng.normal(size=tuple(od.nlg))
    a = v.copy()
    O.lib().orc_update_bcs(C.byref(od), O.ptr(a), C.c_double(bc))
    b = v.copy()
    R.update_all_bcs(rd, b, bc)
    assert np.array_equal(a, b)
