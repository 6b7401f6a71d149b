"""The BOMEX case physics on the GPU against the oracle: the surface fluxes,
Rayleigh damping and large-scale forcing operators on their own, and whole
steps of the reference's Bomex setup (Smagorinsky SGS, Rayleigh damping,
surface and forcing, CFL-limited dt up to 10 s) from InitBomex's state.

Tolerances: the surface entropy flux and the large-scale entropy tendency
evaluate log/pow (device libm against glibc: last bit), the damping and the
geostrophic term use level means summed in a different order, so the
operators are compared within 1e-12 of each tendency's scale; whole FP64
steps within 1e-12."""
import copy

import numpy as np
import pytest
import torch

from tests.helpers import floor_tolerance, oracle_sim, rel_err, ulp_perturbed

pytestmark = pytest.mark.gpu


def _bomex(nx=32, ny=16, nz=75, precision="f64", seed=61, dt=10.0):
    from pycles_amd import cases
    from pycles_amd.Grid import Grid
    from pycles_amd.ReferenceState import ReferenceState
    from pycles_amd.ThermodynamicsSA import ThermodynamicsFactory
    nl = cases.bomex_namelist(nx, ny, nz, dt=dt, precision=precision, seed=seed)
    Gr = Grid(nl)
    RS = ReferenceState(Gr)
    Th = ThermodynamicsFactory(nl)
    fields = cases.init_bomex(Gr, RS, Th, seed=seed)
    # a little resolved motion so every term has something to act on
    rng = np.random.default_rng(seed)
    d = Gr.dims
    for n, amp in (("u", 0.5), ("v", 0.5), ("w", 0.2)):
        f = fields[n].copy()
        f[d.gw:d.gw + d.n[2]] += amp * rng.standard_normal((d.n[2], d.n[1], d.n[0]))
        fields[n] = cases.fill_z_ghosts(f, d.gw, -1.0 if n == "w" else 1.0)
    # divergence-free, as the stepping assumes (a divergent start gives the
    # first stage flux-form scalar tendencies of order s div(u))
    fields["u"], fields["v"], fields["w"] = cases.project_initial(Gr, RS, fields["u"], fields["v"], fields["w"])
    return nl, Gr, RS, Th, fields


def _oracle(Gr, RS, Th, fields, nl, **kw):
    bomex = {"zl": Gr.zl, "zl_half": Gr.zl_half, "u0": RS.u0, "v0": RS.v0,
             "gamma_r": nl["damping"]["Rayleigh"]["gamma_r"], "z_d": nl["damping"]["Rayleigh"]["z_d"]}
    return oracle_sim(Gr, RS, Th, fields, True, sgs={"cs": 0.17, "prt": 1.0 / 3.0}, bomex=bomex, **kw)


def _sim(nl, Gr, RS, Th, fields):
    from pycles_amd.Simulation3d import Simulation3d
    return Simulation3d().initialize(nl, Gr=Gr, Ref=RS, Th=Th, fields=fields)


def _close(g, o, tol=1e-12, scale=0.0):
    scale = max(np.abs(o).max(), scale)
    return np.abs(g - o).max() <= tol * max(scale, 1e-300)


def test_bomex_operators(gpu):
    nl, Gr, RS, Th, fields = _bomex()
    sim = _sim(nl, Gr, RS, Th, fields)
    sim.Th.max_iter = 0
    sim.Th.update(sim.Gr, sim.Ref, sim.PV, sim.DV)
    osim = _oracle(Gr, RS, Th, fields, nl)
    osim.thermo_update()
    gw = Gr.dims.gw
    sl = (slice(gw, -gw),) * 3
    # the same thermodynamic state on both sides
    for n in ("temperature", "qv", "ql"):
        osim.dv[n][...] = sim.DV.get_variable_array(n, Gr)
    osim.pv("qt")[...] = sim.PV.get_variable_array("qt", Gr)
    # the profiles the device uses are the oracle's
    assert np.array_equal(sim.Fo.scheme.profiles_host.reshape(-1), osim.forcing_prof)
    assert np.array_equal(sim.Damping.scheme.gamma_z_host, osim.gamma_z)
    assert np.array_equal(sim.Damping.scheme.gamma_zhalf_host, osim.gamma_zhalf)
    gamma_max = max(np.abs(osim.gamma_z).max(), np.abs(osim.gamma_zhalf).max())
    assert gamma_max == pytest.approx(nl["damping"]["Rayleigh"]["gamma_r"], rel=0.05)
    for step, (gpu_op, orc_op) in enumerate((
            (lambda: sim.Sur.update(sim.Gr, sim.Ref, sim.PV, sim.DV, sim.Pa, sim.TS), osim.surface_update),
            (lambda: sim.Damping.update(sim.Gr, sim.Ref, sim.PV, sim.DV, sim.Pa), osim.damping_update),
            (lambda: sim.Fo.update(sim.Gr, sim.Ref, sim.PV, sim.DV, sim.Pa), osim.forcing_update))):
        sim.PV.zero_tendencies()
        osim.tend[...] = 0.0
        gpu_op()
        orc_op()
        torch.cuda.synchronize()
        for n in osim.pv_names:
            g = sim.PV.get_tendency_array(n, Gr)[sl]
            o = osim.pt(n)[sl]
            # the damping tendency gamma (phi - <phi>) is the difference of two
            # terms of size gamma |phi|; above the noise layer phi is uniform
            # along each level and the tendency is the rounding of the level
            # mean (the device sums in a tree, the reference serially), so it
            # is compared at the scale of those terms
            scale = gamma_max * np.abs(osim.pv(n)[sl]).max() if step == 1 else 0.0
            assert _close(g, o, scale=scale), (step, n, np.abs(g - o).max(), np.abs(o).max(), scale)
    # the surface operator touched the first level, forcing all, damping the top
    fl = sim.Sur.flux2d.cpu().numpy()
    of = osim.surface_flux2d.reshape(8, Gr.dims.nlg[0], Gr.dims.nlg[1])[:, gw:-gw, gw:-gw].transpose(0, 2, 1)
    for r in range(8):
        assert _close(fl[r], of[r]), r


def test_bomex_steps_against_the_oracle(gpu):
    nl, Gr, RS, Th, fields = _bomex(seed=62)
    sim = _sim(nl, Gr, RS, Th, fields)
    sim.Th.max_iter = 0
    sim.run(max_steps=2)
    torch.cuda.synchronize()
    runs = {}
    for key, start in (("oracle", fields), ("oracle_ulp", ulp_perturbed(fields))):
        o = _oracle(Gr, RS, Th, start, nl, dt=10.0, dt_max=10.0, cfl_limit=0.7)
        o.begin_run()
        o.step()
        o.step()
        runs[key] = o
    osim = runs["oracle"]

    def errs_of(get_pv, get_dv):
        e = {n: rel_err(get_pv(n), Gr.to_device_layout(osim.pv(n))) for n in osim.pv_names}
        for n in ("temperature", "ql", "viscosity", "dynamic_pressure"):
            e[n] = rel_err(get_dv(n), Gr.to_device_layout(osim.dv[n]))
        return e
    errs = errs_of(sim.PV.get_device_layout, sim.DV.get_device_layout)
    o2 = runs["oracle_ulp"]
    floor = errs_of(lambda n: Gr.to_device_layout(o2.pv(n)), lambda n: Gr.to_device_layout(o2.dv[n]))
    tol = floor_tolerance(floor)
    print("2 FP64 BOMEX steps vs oracle", errs, "floor (1 ulp of s)", floor, "t", sim.TS.t, osim.t,
          "dt", sim.TS.dt, osim.dt)
    assert sim.TS.t == osim.t and sim.TS.dt == pytest.approx(osim.dt, rel=1e-12)
    assert all(errs[n] <= tol[n] for n in errs), (errs, tol)


def test_bomex_fp32_tracks_fp64(gpu):
    out = {}
    for precision in ("f64", "f32"):
        nl, Gr, RS, Th, fields = _bomex(32, 32, 75, precision=precision, seed=63)
        sim = _sim(nl, Gr, RS, Th, fields)
        sim.run(max_steps=3)
        torch.cuda.synchronize()
        out[precision] = sim
    gw = out["f64"].Gr.dims.gw
    for n in ("s", "qt"):
        a = out["f32"].PV.get_device_layout(n)[gw:-gw]
        b = out["f64"].PV.get_device_layout(n)[gw:-gw]
        assert np.max(np.abs(a - b) / np.abs(b)) < 1e-5, n
    a = out["f32"].DV.get_device_layout("temperature")[gw:-gw]
    b = out["f64"].DV.get_device_layout("temperature")[gw:-gw]
    assert np.max(np.abs(a - b) / b) < 1e-5
