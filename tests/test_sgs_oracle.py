"""The oracle's kinematics, Smagorinsky SGS and diffusion restatements.

Bitwise against the reference's own compiled kinematics.h and
momentum_diffusion.h (oracle/_ref, built from the reference tree where it is
present).  The Smagorinsky closure (sgs.h) and the scalar diffusion
(scalar_diffusion.h) include the generated parameters.h and are not built;
they are pinned by what the mathematics fixes: the strain-rate magnitude of
a uniform shear, the Richardson-number cut-off Ri = N^2/|S|^2 >= Pr of the
stratified closure (Lilly 1962), the exactness of the second-order diffusion
on quadratic profiles with the reference's density weights, and the
conservation of the mass-weighted scalar.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from oracle import reflib

needs_ref = pytest.mark.skipif(not reflib.available(), reason="reference tree not present")


def _sim(n=(7, 6, 8), gw=3, dx=(40.0, 30.0, 20.0), seed=0, sgs=None):
    rng = np.random.default_rng(seed)
    d = O.make_dims(*n, gw, *dx)
    n2 = d.nlg[2]
    rho0 = rng.uniform(0.8, 1.2, n2)
    rho0h = rng.uniform(0.8, 1.2, n2)
    ref = O.Ref(np.full(n2, 1e5), np.full(n2, 1e5), 1.0 / rho0, 1.0 / rho0h, rho0, rho0h)
    sim = O.OracleSimulation(d, ref, True, None, sgs=sgs or {"cs": 0.17, "prt": 1.0 / 3.0})
    return sim, d, rng


def _random_velocities(sim, d, rng):
    for name, scale in (("u", 3.0), ("v", 2.0), ("w", 1.0)):
        sim.pv(name)[...] = rng.normal(size=tuple(d.nlg)) * scale
    sim.update_all_bcs()


@needs_ref
def test_kinematics_bitwise_against_reference():
    sim, d, rng = _sim(seed=1)
    _random_velocities(sim, d, rng)
    sim.kinematics_update()
    rd = reflib.make_dims(*d.n, d.gw, *d.dx)
    vgrad, strain, smag = reflib.kinematics(rd, [sim.pv(n) for n in ("u", "v", "w")])
    assert np.array_equal(sim.vgrad.reshape(-1), vgrad)
    assert np.array_equal(sim.strain.reshape(-1), strain)
    assert np.array_equal(sim.smag.reshape(-1), smag)


@needs_ref
def test_momentum_diffusion_bitwise_against_reference():
    sim, d, rng = _sim(seed=2)
    _random_velocities(sim, d, rng)
    sim.kinematics_update()
    sim.dv["viscosity"][...] = rng.uniform(0.0, 5.0, size=tuple(d.nlg))
    sim.dv["temperature"][...] = rng.uniform(280.0, 300.0, size=tuple(d.nlg))
    for n in ("u", "v", "w", "s"):
        sim.pt(n)[...] = rng.normal(size=tuple(d.nlg))
    want = [sim.pt(n).copy() for n in ("u", "v", "w")]
    want_s = sim.pt("s").copy()
    sim.momentum_diffusion_update()
    rd = reflib.make_dims(*d.n, d.gw, *d.dx)
    r = sim.ref.arrays  # p0, p0_half, alpha0, alpha0_half, rho0, rho0_half
    reflib.momentum_diffusion(rd, r[4], r[5], r[2], r[3], sim.strain.reshape(-1), sim.dv["viscosity"], want,
                              sim.smag, sim.dv["temperature"], want_s)
    for q, n in enumerate(("u", "v", "w")):
        assert np.array_equal(sim.pt(n), want[q]), n
    assert np.array_equal(sim.pt("s"), want_s)


def test_strain_rate_magnitude_of_uniform_shear():
    """u = a z: the only strain components are S_13 = S_31 = a/2, so
    |S| = sqrt(2 S_ij S_ij) = |a| at every point the reference computes it
    ([1, nlg-1) in each direction).  u = b x gives S_11 = b; the reference's
    off-diagonal loop (kinematics.h:111-128) starts its second index at the
    first, so S_11 enters once more with weight 2: |S| = sqrt(2 (b^2 + 2 b^2))
    = sqrt(6)|b| where all four interpolation points exist (i >= 2)."""
    sim, d, rng = _sim(seed=3)
    a = 0.013
    k = np.arange(d.nlg[2], dtype=np.float64)
    sim.pv("u")[...] = a * k[None, None, :] * d.dx[2]
    sim.kinematics_update()
    inner = sim.smag[1:-1, 1:-1, 1:-1]
    assert np.allclose(inner, abs(a), rtol=1e-13, atol=0)
    assert np.all(sim.smag[0] == 0.0) and np.all(sim.smag[:, :, -1] == 0.0)
    sim.pv("u")[...] = 0.0
    b = -0.02
    i = np.arange(d.nlg[0], dtype=np.float64)
    sim.pv("u")[...] = b * i[:, None, None] * d.dx[0]
    sim.kinematics_update()
    assert np.allclose(sim.smag[2:-1, 1:-1, 1:-1], np.sqrt(6.0) * abs(b), rtol=1e-13, atol=0)


def test_smagorinsky_closure():
    """nu = (cs Delta)^2 |S| with Delta = (dx dy dz)^(1/3) in neutral or
    unstable air; in stable air the factor sqrt(1 - Ri/Pr) with Ri = N^2/|S|^2,
    zero from Ri >= Pr on; diffusivity = nu/Pr."""
    sim, d, rng = _sim(dx=(20.0, 20.0, 20.0), seed=4)
    cs, prt = sim.sgs["cs"], sim.sgs["prt"]
    delta = 20.0
    S = rng.uniform(0.01, 0.1, size=tuple(d.nlg))
    sim.smag[...] = S
    ri = np.linspace(-1.0, 1.0, d.nlg[2])[None, None, :] * np.ones(tuple(d.nlg))
    sim.dv["buoyancy_frequency"][...] = ri * S * S
    sim.sgs_update()
    nu = sim.dv["viscosity"]
    neutral = (cs * delta) ** 2 * S
    fac = np.where(ri > 0, np.sqrt(np.maximum(1.0 - ri / prt, 0.0)), 1.0)
    assert np.allclose(nu, neutral * fac, rtol=1e-12, atol=1e-15)
    assert np.all(nu[ri >= prt] == 0.0) and np.all(nu[ri < prt] > 0.0)
    assert np.allclose(sim.dv["diffusivity"], nu / prt, rtol=1e-15, atol=0)


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_scalar_diffusion_of_a_quadratic(direction):
    """phi = c x^2 with a uniform diffusivity K: the second-order flux
    -K (phi[i+1] - phi[i])/dx * rho is exact on the face, so the tendency is
    -a0h[k] (F[i] - F[i-1])/dx with F = -K c (2 x_face) rho: in x and y
    K * 2c * rho0_half * a0h = 2 K c; in z with the face densities rho0[k]."""
    sim, d, rng = _sim(seed=5 + direction)
    K = 3.5
    c = 0.7
    h = d.dx[direction]
    idx = np.indices(tuple(d.nlg), dtype=np.float64)[direction]
    x = (idx - d.gw) * h / 100.0
    sim.pv("s")[...] = c * x * x
    sim.pv("qt")[...] = 0.0
    sim.dv["diffusivity"][...] = K
    sim.scalar_diffusion_update()
    gw = d.gw
    sl = (slice(gw, -gw),) * 3
    r = sim.ref.arrays
    rho0, rho0h, a0h = r[4], r[5], r[3]
    hs = h / 100.0
    flux = lambda xf, rho: -K * (2.0 * c * xf * hs) * rho / h  # noqa: E731  (phi[i+1]-phi[i] = 2c x_face hs)
    if direction == 2:
        up = flux(x + 0.5 * hs, rho0[None, None, :])
        dn = flux(x - 0.5 * hs, np.roll(rho0, 1)[None, None, :])
    else:
        up = flux(x + 0.5 * hs, rho0h[None, None, :])
        dn = flux(x - 0.5 * hs, rho0h[None, None, :])
    want = -a0h[None, None, :] * (up - dn) / h
    if direction == 2:
        # the surface flux (face below the first interior level) is zero
        want[:, :, gw] = -a0h[gw] * (up[:, :, gw] - 0.0) / h
    got = sim.pt("s")[sl]
    assert np.allclose(got, want[sl], rtol=1e-11, atol=1e-14 * np.abs(want[sl]).max())


def test_scalar_diffusion_conserves():
    """Periodic x, y, the zero surface flux and a mirrored top: the
    mass-weighted sum of the diffusive tendency vanishes."""
    sim, d, rng = _sim(seed=9)
    sim.pv("s")[...] = rng.normal(size=tuple(d.nlg))
    sim.pv("qt")[...] = rng.normal(size=tuple(d.nlg))
    sim.update_all_bcs()
    sim.dv["diffusivity"][...] = rng.uniform(0.5, 2.0, size=tuple(d.nlg))
    O.lib().orc_update_bcs(C.byref(d), O.ptr(sim.dv["diffusivity"]), C.c_double(1.0))
    sim.scalar_diffusion_update()
    gw = d.gw
    rho0h = sim.ref.arrays[5]
    for n in ("s", "qt"):
        m = sim.pt(n)[gw:-gw, gw:-gw, gw:-gw] * rho0h[gw:-gw][None, None, :]
        assert abs(m.sum()) <= 1e-13 * np.abs(m).sum(), n
