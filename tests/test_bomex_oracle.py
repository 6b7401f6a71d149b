"""Pins of the oracle's BOMEX case physics (Surface.pyx:297-373,
Forcing.pyx:102-203, 1982-2040, Damping.pyx:54-133) by what the physics and
mathematics fix:

- the surface entropy flux is the entropy change per unit of the theta and
  qt fluxes (a finite difference of the entropy function at the surface
  state), and the momentum flux is -u*^2 u/|U| (a uniform wind gives
  -u*^2 times its direction cosine);
- the geostrophic term vanishes when the mean wind equals the geostrophic
  wind, and is f (u_g - <u>) and -f (v_g - <v>) otherwise;
- the large-scale entropy tendency of a temperature tendency at fixed qt is
  cpm dT/T (the entropy's temperature derivative) and of a moisture
  tendency (s_v - s_d) dqt;
- the subsidence of a field linear in z is -w_s(z) times its gradient, with
  the top interior level repeating the level below, as the reference;
- the Rayleigh coefficient is gamma_r sin^2 of the depth into the layer,
  gamma_r at the top and 0 at its bottom, and damping removes gamma times
  the deviation from the level mean."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from pycles_amd import parameters as P


def _state(n=(6, 5, 40), gw=3, dz=40.0, seed=0):
    rng = np.random.default_rng(seed)
    d = O.make_dims(*n, gw, 100.0, 100.0, dz)
    n2 = d.nlg[2]
    zl = (np.arange(n2) - gw + 1.0) * dz
    zl_half = (np.arange(n2) - gw + 0.5) * dz
    p0 = 1.0e5 * np.exp(-zl_half / 8000.0)
    a0 = rng.uniform(0.8, 1.2, n2)
    a0h = rng.uniform(0.8, 1.2, n2)
    ref = O.Ref(p0, p0, a0, a0h, 1.0 / a0, 1.0 / a0h)
    bomex = {"zl": zl, "zl_half": zl_half, "u0": -6.59, "v0": 0.0, "gamma_r": 0.2, "z_d": 600.0}
    sim = O.OracleSimulation(d, ref, True, None, bomex=bomex)
    return sim, d, rng


def test_surface_fluxes():
    sim, d, rng = _state()
    L = O.lib()
    gw = d.gw
    T, qt, p0 = 299.0, 0.017, sim.ref.arrays[1][gw]
    sim.dv["temperature"][...] = T
    sim.pv("qt")[...] = qt
    sim.dv["qv"][...] = qt
    sim.dv["ql"][...] = 0.0
    U = 3.0
    sim.pv("u")[...] = U - sim.bomex["u0"]
    sim.pv("v")[...] = 0.0
    sim.surface_update()
    tf = sim.ref.arrays[3][gw] / sim.ref.arrays[2][gw - 1] / d.dx[2]
    # entropy flux = ds per unit theta and qt flux at the surface state
    h = 1e-3
    theta = T / L.orc_exner(p0)
    s0 = L.orc_entropy_sa(p0, T, qt, 0.0, 0.0)
    s1 = L.orc_entropy_sa(p0, (theta + h * 8.0e-3) * L.orc_exner(p0), qt + h * 5.2e-5, 0.0, 0.0)
    sflux = L.orc_entropyflux_from_thetaflux_qtflux(8.0e-3, 5.2e-5, p0, T, qt, qt)
    assert sflux == pytest.approx((s1 - s0) / h, rel=1e-5)
    sl = (slice(gw, -gw), slice(gw, -gw), gw)
    assert np.allclose(sim.pt("s")[sl], sflux * tf, rtol=1e-14, atol=0)
    assert np.allclose(sim.pt("qt")[sl], 5.2e-5 * tf, rtol=1e-14, atol=0)
    # uniform wind along x: tau_x = -u*^2 (u/|U|) |U|/|U| = -u*^2, tau_y = 0
    assert np.allclose(sim.pt("u")[sl], -(0.28 ** 2) * tf, rtol=1e-14, atol=0)
    assert np.all(sim.pt("v")[sl] == 0.0)
    # nothing above the first level
    assert np.all(sim.pt("s")[gw:-gw, gw:-gw, gw + 1:] == 0.0)


def test_geostrophic_term_and_large_scale_tendencies():
    sim, d, rng = _state(seed=1)
    gw = d.gw
    prof = sim.forcing_prof.reshape(5, -1)
    ug, vg, dtdt, dqtdt, sub = prof
    u0, v0 = sim.bomex["u0"], sim.bomex["v0"]
    # in geostrophic balance and without vertical structure, only subsidence
    # of nothing and the large-scale sources act
    sim.pv("u")[...] = (ug - u0)[None, None, :]
    sim.pv("v")[...] = (vg - v0)[None, None, :]
    T, qt = 295.0, 0.012
    sim.dv["temperature"][...] = T
    sim.dv["ql"][...] = 0.0
    sim.pv("qt")[...] = qt
    sim.pv("s")[...] = 7000.0
    sim.forcing_update()
    k = np.arange(gw, d.nlg[2] - gw)
    sl = (slice(gw, -gw), slice(gw, -gw), k)
    # in balance the Coriolis term leaves only the rounding of (u_g - u0) + u0;
    # v_g = 0 is uniform, so v has no subsidence either
    assert np.abs(sim.pt("v")[sl]).max() <= 1e-15 * 0.376e-4 * 20.0
    p0 = sim.ref.arrays[1]
    L = O.lib()
    cpm = (1 - qt) * P.cpd + qt * P.cpv
    # dq/dt from the entropy derivatives of a state at fixed p and T
    e = 1e-7
    dsdq = np.array([(L.orc_entropy_sa(p0[kk], T, qt + e, 0.0, 0.0) - L.orc_entropy_sa(p0[kk], T, qt - e, 0.0, 0.0))
                     / (2 * e) for kk in k])
    want_s = cpm * dtdt[k] / T + dsdq * dqtdt[k]
    got_s = sim.pt("s")[gw + 1, gw + 1, k]
    assert np.allclose(got_s, want_s, rtol=1e-5, atol=1e-12)
    assert np.allclose(sim.pt("qt")[gw + 1, gw + 1, k], dqtdt[k], rtol=0, atol=0)
    # Coriolis acts on the departure from the geostrophic wind
    sim2, d2, _ = _state(seed=2)
    sim2.pv("u")[...] = 1.5
    sim2.pv("v")[...] = -0.5
    sim2.dv["temperature"][...] = T
    sim2.pv("qt")[...] = qt
    sim2.forcing_update()
    f = 0.376e-4
    ug2, vg2 = sim2.forcing_prof.reshape(5, -1)[:2]
    assert np.allclose(sim2.pt("u")[gw + 2, gw, k], -f * (vg2[k] - (-0.5 + v0)), rtol=1e-13, atol=0)
    assert np.allclose(sim2.pt("v")[gw + 2, gw, k], f * (ug2[k] - (1.5 + u0)), rtol=1e-13, atol=0)


def test_subsidence_of_a_linear_profile():
    sim, d, rng = _state(seed=3)
    gw = d.gw
    sub = sim.forcing_prof.reshape(5, -1)[4]
    a = 0.004
    z = sim.bomex["zl_half"]
    sim.pv("u")[...] = 0.0
    sim.pv("v")[...] = 0.0
    sim.pv("s")[...] = 7000.0 + a * z[None, None, :]
    sim.pv("qt")[...] = 0.01
    sim.dv["temperature"][...] = 290.0
    sim.forcing_update()
    tend = sim.pt("s")[gw + 1, gw + 1]
    dtdt, dqtdt = sim.forcing_prof.reshape(5, -1)[2:4]
    ls = (1 - 0.01) * P.cpd + 0.01 * P.cpv
    k = np.arange(gw, d.nlg[2] - gw - 1)
    ls_part = ls * dtdt[k] / 290.0
    # qt is uniform: the drying source on s is (sv - sd) dqtdt, subtract it via the qt-only case
    sim_q, _, _ = _state(seed=3)
    sim_q.pv("s")[...] = 7000.0
    sim_q.pv("qt")[...] = 0.01
    sim_q.dv["temperature"][...] = 290.0
    sim_q.forcing_update()
    sub_part = tend[k] - sim_q.pt("s")[gw + 1, gw + 1, k]
    assert np.allclose(sub_part, -a * sub[k], rtol=1e-9, atol=1e-15)
    top = d.nlg[2] - gw - 1
    assert tend[top] - sim_q.pt("s")[gw + 1, gw + 1, top] == pytest.approx(-a * sub[top - 1], rel=1e-9)
    del ls_part


def test_rayleigh_profile_and_damping():
    sim, d, rng = _state(seed=4)
    gw = d.gw
    z_top = d.dx[2] * d.n[2]
    zl, zh = sim.bomex["zl"], sim.bomex["zl_half"]
    g, gh = sim.gamma_z, sim.gamma_zhalf
    assert np.all(g[zl < z_top - 600.0] == 0.0) and np.all(gh[zh < z_top - 600.0] == 0.0)
    top = np.argmin(np.abs(zl - z_top))
    assert g[top] == pytest.approx(0.2, rel=1e-15)
    sel = zh >= z_top - 600.0
    assert np.allclose(gh[sel], 0.2 * np.sin(np.pi / 2 * (1 - (z_top - zh[sel]) / 600.0)) ** 2, rtol=1e-14)
    for n in sim.pv_names:
        sim.pv(n)[...] = rng.normal(size=tuple(d.nlg))
    sim.damping_update()
    sl = slice(gw, -gw)
    for n in sim.pv_names:
        v = sim.pv(n)
        mean = v[sl, sl, :].mean(axis=(0, 1))
        gam = gh if n == "w" else g
        want = -(v - mean[None, None, :]) * gam[None, None, :]
        got = sim.pt(n)
        assert np.allclose(got[sl, sl, sl], want[sl, sl, sl], rtol=1e-12, atol=1e-15), n
