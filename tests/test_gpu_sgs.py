"""Kinematics, Smagorinsky SGS and diffusion on the GPU against the oracle
(which is bitwise equal to the reference's own kinematics.h and
momentum_diffusion.h, and pinned by mathematics for sgs.h and
scalar_diffusion.h).

FP64 tolerances: the SGS kernel and the diffusion kernel use + - * / and
sqrt (correctly rounded on both sides) and a cbrt evaluated on the host, so
|S|, nu, D and the diffusive tendencies are compared bitwise; whole steps
include the thermodynamics' transcendentals and are compared at 1e-12."""
import copy

import numpy as np
import pytest
import torch

from tests.helpers import floor_tolerance, host_case, oracle_sim, rel_err, ulp_perturbed

pytestmark = pytest.mark.gpu

SGS_NL = {"scheme": "Smagorinsky", "Smagorinsky": {"cs": 0.17, "prt": 1.0 / 3.0}}
SGS_OR = {"cs": 0.17, "prt": 1.0 / 3.0}


def _case(nx, ny, nz, seed, precision="f64", gw=3, order=(5, 5), rough=True):
    from tests.test_gpu_parity import _saturating, _swirl
    from pycles_amd import cases
    nl, Gr, RS, Th, fields = host_case(nx, ny, nz, gw=gw, seed=seed, order=order)
    fields = _swirl(Gr, fields, amp=3.0)
    if rough:
        fields = _saturating(Gr, fields)
    # a divergent start makes the first stage's flux-form scalar tendencies
    # s div(u) (hundreds of J/kg/K/s) and drives T off the lookup table
    fields["u"], fields["v"], fields["w"] = cases.project_initial(Gr, RS, fields["u"], fields["v"], fields["w"])
    nl = copy.deepcopy(nl)
    nl["sgs"] = copy.deepcopy(SGS_NL)
    nl["mi355x"]["precision"] = precision
    return nl, Gr, RS, Th, fields


def _sim(nl, Gr, RS, Th, fields):
    from pycles_amd.Simulation3d import Simulation3d
    return Simulation3d().initialize(nl, Gr=Gr, Ref=RS, Th=Th, fields=fields)


def test_sgs_and_diffusion_operators_bitwise(gpu):
    nl, Gr, RS, Th, fields = _case(24, 16, 20, seed=51)
    sim = _sim(nl, Gr, RS, Th, fields)
    sim.Th.max_iter = 0
    sim.Th.update(sim.Gr, sim.Ref, sim.PV, sim.DV)
    osim = oracle_sim(Gr, RS, Th, fields, True, sgs=SGS_OR)
    osim.thermo_update()
    # the GPU's N^2 where the oracle's is defined, so both closures see the same input
    gw = Gr.dims.gw
    sim.Ke.update(sim.Gr, sim.PV)
    sim.SGS.update(sim.Gr, sim.DV, sim.PV, sim.Ke)
    osim.dv["buoyancy_frequency"][...] = sim.DV.get_variable_array("buoyancy_frequency", Gr)
    osim.kinematics_update()
    osim.sgs_update()
    torch.cuda.synchronize()
    lv = slice(1, Gr.dims.nzg - 1)
    smag = sim.Ke.strain_rate_mag.to(torch.float64).cpu().numpy()
    assert np.array_equal(smag[lv], Gr.to_device_layout(osim.smag)[lv])
    for n in ("viscosity", "diffusivity"):
        assert np.array_equal(sim.DV.get_device_layout(n)[lv], Gr.to_device_layout(osim.dv[n])[lv]), n
    # diffusion on top of the same (random) tendencies
    rng = np.random.default_rng(5)
    for n in osim.pv_names:
        t = rng.normal(size=osim.pt(n).shape)
        osim.pt(n)[...] = t
        sim.PV.tend(n).copy_(torch.from_numpy(Gr.to_device_layout(t)))
    osim.dv["viscosity"][...] = sim.DV.get_variable_array("viscosity", Gr)
    osim.dv["diffusivity"][...] = sim.DV.get_variable_array("diffusivity", Gr)
    osim.dv["temperature"][...] = sim.DV.get_variable_array("temperature", Gr)
    osim.smag[...] = Gr.from_device_layout(smag)
    from pycles_amd.SGS import diffuse
    diffuse(sim.SD, sim.MD, sim.Gr, sim.Ref, sim.PV, sim.DV, sim.Ke)
    osim.scalar_diffusion_update()
    osim.momentum_diffusion_update()
    torch.cuda.synchronize()
    for n in osim.pv_names:
        g = sim.PV.get_tendency_array(n, Gr)[gw:-gw, gw:-gw, gw:-gw]
        o = osim.pt(n)[gw:-gw, gw:-gw, gw:-gw]
        assert np.array_equal(g, o), f"{n}: max diff {np.abs(g - o).max()} (rel {rel_err(g, o)})"


def test_steps_with_sgs_match_the_oracle(gpu):
    nl, Gr, RS, Th, fields = _case(32, 16, 24, seed=52)
    sim = _sim(nl, Gr, RS, Th, fields)
    sim.Th.max_iter = 0
    sim.run(max_steps=2)
    torch.cuda.synchronize()
    runs = {}
    for key, start in (("oracle", fields), ("oracle_ulp", ulp_perturbed(fields))):
        osim = oracle_sim(Gr, RS, Th, start, True, sgs=SGS_OR)
        osim.begin_run()
        osim.step()
        osim.step()
        runs[key] = osim
    osim = runs["oracle"]

    def errs_of(get_pv, get_dv):
        e = {n: rel_err(get_pv(n), Gr.to_device_layout(osim.pv(n))) for n in osim.pv_names}
        for n in ("temperature", "viscosity", "dynamic_pressure"):
            e[n] = rel_err(get_dv(n), Gr.to_device_layout(osim.dv[n]))
        return e
    errs = errs_of(sim.PV.get_device_layout, sim.DV.get_device_layout)
    o2 = runs["oracle_ulp"]
    floor = errs_of(lambda n: Gr.to_device_layout(o2.pv(n)), lambda n: Gr.to_device_layout(o2.dv[n]))
    tol = floor_tolerance(floor)
    print("2 FP64 steps with Smagorinsky SGS and diffusion vs oracle", errs, "floor (1 ulp of s)", floor)
    assert all(errs[n] <= tol[n] for n in errs), (errs, tol)
    # the closure is active: the viscosity is not negligible
    assert float(sim.DV.field("viscosity").abs().max()) > 1e-2


def test_fp32_with_sgs_tracks_fp64(gpu):
    """FP32 against FP64 over 3 steps with the closure on, from a smooth start
    (the seeded moist state with a swirl), gated at 1e-5 or at twice what
    FP32 storage alone does to an FP64 run (its state rounded to FP32 storage
    after every step).  The rough start of the FP64 tests (entropy noise of
    25 J/kg/K cell to cell) is not used here: its Richardson numbers sit at
    the Smagorinsky switch sqrt(max(1 - N^2/(Pr S^2), 0)) in many cells, where
    the FP32 saturation adjustment's 3e-4 K tolerance alone moves nu by tens
    of percent within a step."""
    import bench
    out = {}
    for precision in ("f64", "f32", "storage"):
        nl, Gr, RS, Th, fields = _case(32, 32, 32, seed=53, precision="f64" if precision == "storage" else precision,
                                       rough=False)
        start = bench.perturbed_fields(fields, "fp32") if precision == "storage" else fields
        sim = _sim(nl, Gr, RS, Th, start)
        rounding = bench.FP32Storage(sim, fields) if precision == "storage" else None
        sim.begin_run()
        for _ in range(3):
            sim.step()
            if rounding is not None:
                rounding.round()
        torch.cuda.synchronize()
        out[precision] = sim
    gw = out["f64"].Gr.dims.gw

    def rel(n, which, dv=False):
        src = (lambda s: s.DV) if dv else (lambda s: s.PV)
        a = src(out[which]).get_device_layout(n)[gw:-gw]
        b = src(out["f64"]).get_device_layout(n)[gw:-gw]
        return float(np.max(np.abs(a - b) / np.abs(b)))
    for n, dv in (("s", False), ("qt", False), ("temperature", True)):
        err, floor = rel(n, "f32", dv), rel(n, "storage", dv)
        print(f"{n}: FP32 {err:.3e}, FP64 rounded to FP32 storage every step {floor:.3e}", flush=True)
        assert err <= max(1e-5, 2.0 * floor), (n, err, floor)
    assert float(out["f32"].DV.field("viscosity").abs().max()) > 1e-2  # the closure is active
    for n in ("u", "v", "w"):
        a = out["f32"].PV.get_device_layout(n)
        b = out["f64"].PV.get_device_layout(n)
        c = out["storage"].PV.get_device_layout(n)
        scale = max(1.0, np.max(np.abs(b)))
        assert np.max(np.abs(a - b)) < max(1e-3 * scale, 2.0 * np.max(np.abs(c - b))), n
