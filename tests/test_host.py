"""Host-side logic on the CPU: the shared thermodynamic inputs (golden
fixture, hydrostatic balance, uniform entropy of the reference state), the
layout conversions, the counter-based initial-noise generator and the
divergence-free initial velocities."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from pycles_amd import cases
from pycles_amd import parameters as P
from pycles_amd.Grid import Grid
from pycles_amd.ReferenceState import ReferenceState
from pycles_amd.ThermodynamicsSA import ThermodynamicsFactory
from tests.helpers import host_case

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "thermo_profiles.npz")


def test_thermo_inputs_match_golden():
    from tests.golden.make_golden import compute
    want = np.load(GOLDEN)
    got = compute()
    assert set(want.files) == set(got)
    for k in want.files:
        np.testing.assert_allclose(got[k], want[k], rtol=1e-12, atol=0, err_msg=k)


@pytest.mark.parametrize("moist", [True, False])
def test_reference_state_is_hydrostatic_and_isentropic(moist):
    """d ln p0/dz = -g / (Rd T (1 - qt + qv/eps_v)) (ReferenceState.pyx:54-56)
    and the saturation-adjusted entropy is uniform (:112-121, tightened)."""
    nl, Gr, RS, Th, _ = host_case(4, 4, 64, moist=moist, dz=25.0)
    gw = Gr.dims.gw
    z = Gr.z_half[gw:-gw]
    lp = np.log(RS.p0_half[gw:-gw])
    dlp = np.gradient(lp, z)
    lt = O.Lookup(Th.CC.T, Th.CC.pv) if moist else None
    for k in range(2, len(z) - 2):
        p0 = RS.p0_half[gw + k]
        if moist:
            T, qv, ql, qi, _ = O.eos_sa(lt, p0, RS.sg, RS.qtg)
            s = O.lib().orc_entropy_sa(p0, T, RS.qtg, ql, qi)
            assert abs(s - RS.sg) / RS.sg < 1e-5
        else:
            T = O.lib().orc_eos_dry(p0, RS.sg)
            ql = qi = 0.0
        rhs = -P.g / (P.Rd * T * (1.0 - RS.qtg + P.eps_vi * (RS.qtg - ql - qi)))
        assert dlp[k] == pytest.approx(rhs, rel=2e-4)
        # alpha0_half is the specific volume of that state
        qv = RS.qtg - ql - qi
        assert RS.alpha0_half[gw + k] == pytest.approx(O.lib().orc_alpha(p0, T, RS.qtg, qv), rel=1e-10)


@pytest.mark.parametrize("qtg", [0.0, 0.004])
def test_reference_state_against_the_adiabat_in_closed_form(qtg):
    """With a constant entropy and an unsaturated constant qt the reference
    column is an adiabat in closed form: s(p, T, qt) constant with pd, pv
    proportional to p gives T proportional to p^(Rm/cpm), and the hydrostatic
    equation d ln p/dz = -g/(Rm T) then gives T = Tg - g z/cpm and
    p = Pg (T/Tg)^(cpm/Rm), with Rm = (1 - qt) Rd + qt Rv and
    cpm = (1 - qt) cpd + qt cpv.  The reference state (ReferenceState.pyx:
    39-135, odeint with hmax = 1) must match it on full and half levels to
    the integrator's accuracy."""
    from pycles_amd.ThermodynamicsSA import ThermodynamicsSA
    nl = cases.namelist(4, 4, 80, 3, 50.0, 50.0, 25.0, moist=qtg > 0)
    Gr = Grid(nl)
    Th = ThermodynamicsSA(nl) if qtg > 0 else ThermodynamicsFactory(nl)
    RS = ReferenceState(Gr)
    RS.Pg, RS.Tg, RS.qtg = 1.0e5, 300.0, qtg
    RS.initialize(Gr, Th)
    gw = Gr.dims.gw
    Rm = (1.0 - qtg) * P.Rd + qtg * P.Rv
    cpm = (1.0 - qtg) * P.cpd + qtg * P.cpv

    def adiabat(z):
        T = RS.Tg - P.g * z / cpm
        return RS.Pg * (T / RS.Tg) ** (cpm / Rm), T

    # stay below the lifting condensation level in the moist column
    top = 2000.0 if qtg > 0 else 2000.0
    for z_arr, p_arr, a_arr in ((Gr.z_half, RS.p0_half, RS.alpha0_half), (Gr.z, RS.p0, RS.alpha0)):
        sel = [k for k in range(gw, Gr.dims.nlg[2] - gw) if 0.0 <= z_arr[k] <= top]
        assert len(sel) > 40
        for k in sel:
            p, T = adiabat(z_arr[k])
            assert p_arr[k] == pytest.approx(p, rel=2e-8), (k, z_arr[k])
            assert a_arr[k] == pytest.approx(Rm * T / p, rel=2e-8), (k, z_arr[k])


def test_layout_round_trip():
    nl = cases.namelist(6, 5, 4, 3, 1.0, 1.0, 1.0, False)
    Gr = Grid(nl)
    d = Gr.dims
    rng = np.random.default_rng(1)
    dev = rng.normal(size=(d.nzg, d.n[1], d.n[0]))
    ref = Gr.from_device_layout(dev)
    assert ref.shape == tuple(d.nlg)
    assert np.array_equal(Gr.to_device_layout(ref), dev)
    gw = d.gw
    # periodic images in x and y
    assert np.array_equal(ref[:gw], ref[d.n[0]:d.n[0] + gw])
    assert np.array_equal(ref[:, -gw:], ref[:, gw:2 * gw])
    assert ref[gw + 2, gw + 1, 5] == dev[5, 1, 2]


def test_noise_generator():
    shape = (8, 16, 32)
    a = cases.uniform(7, 3, shape)
    assert np.array_equal(a, cases.uniform(7, 3, shape))
    assert a.min() >= 0.0 and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.02 and abs(a.var() - 1.0 / 12.0) < 0.01
    b = cases.uniform(7, 4, shape)
    c = cases.uniform(8, 3, shape)
    assert abs(np.corrcoef(a.ravel(), b.ravel())[0, 1]) < 0.05
    assert abs(np.corrcoef(a.ravel(), c.ravel())[0, 1]) < 0.05
    g = cases.normal(7, 1, (64, 64, 16))
    assert abs(g.mean()) < 0.02 and abs(g.std() - 1.0) < 0.02
    # counter-based: a sub-block is a prefix of the same stream
    assert np.array_equal(cases.uniform(7, 3, (4096,))[:100], cases.uniform(7, 3, (100,)))


@pytest.mark.parametrize("moist", [True, False])
def test_initial_velocities_are_divergence_free(moist):
    nl, Gr, RS, Th, fields = host_case(16, 16, 12, moist=moist, seed=4, sigma=1.0)
    d = Gr.dims
    dims = O.make_dims(d.n[0], d.n[1], d.n[2], d.gw, *d.dx)
    ref = O.Ref(RS.p0, RS.p0_half, RS.alpha0, RS.alpha0_half, RS.rho0, RS.rho0_half)
    sim = O.OracleSimulation(dims, ref, False, None)
    for n in ("u", "v", "w"):
        sim.pv(n)[...] = Gr.from_device_layout(fields[n])
    gw = d.gw
    div = sim.divergence()[gw:-gw, gw:-gw, gw:-gw]
    scale = np.abs(Gr.from_device_layout(fields["u"])).max() / d.dx[0]
    assert np.abs(div).max() < 1e-12 * scale


def test_nan_cfl_stops_the_run():
    """A NaN or negative CFL maximum raises instead of giving a NaN dt (the
    device read itself returns PYCLES_ERR_NONFINITE first on a GPU; the
    GPU test of the boundary covers that)."""
    import math
    from pycles_amd.TimeStepping import TimeStepping
    ts = TimeStepping.__new__(TimeStepping)
    ts.dt, ts.dt_max, ts.cfl_limit = 1.0, 10.0, 0.7
    for bad in (math.nan, -1.0):
        with pytest.raises(RuntimeError, match="killing simulation"):
            ts.accept_cfl(bad)
    ts.accept_cfl(0.35)
    assert ts.cfl_time_step() == 2.0


def test_residual_check_logarithm_method():
    """The FP32 thermodynamics checks its saturated results in FP64 with a
    logarithm of its own (thermo_f32.h, log64): x = m 2^e, m in [sqrt(1/2),
    sqrt(2)), ln x = e ln 2 + 2 atanh(t), t = (m - 1)/(m + 1), the series to
    t^17.  Over the arguments the check meets (T/T~, pd/p~, pv/p~: 6e-6 to 20)
    the truncated series must stay within 1e-14 of the library logarithm,
    far inside the check's 3e-3 J/kg/K on terms of O(100) J/kg/K."""
    import math
    xs = np.exp(np.linspace(np.log(6e-6), np.log(20.0), 20001))
    worst = 0.0
    for x in xs:
        m, e = math.frexp(float(x))
        if m < 0.70710678118654752:
            m *= 2.0
            e -= 1
        t = (m - 1.0) / (m + 1.0)
        t2 = t * t
        p = 1.0 / 17
        for d in (15, 13, 11, 9, 7, 5, 3, 1):
            p = p * t2 + 1.0 / d
        worst = max(worst, abs(e * 0.69314718055994531 + 2.0 * t * p - math.log(x)))
    assert worst < 1e-14
    # the truncation is what bounds it: one term fewer is ~40x worse
    m = math.sqrt(2.0) - 1e-12
    t = (m - 1.0) / (m + 1.0)
    assert t ** 19 / 19 * 2 < 1e-15 < t ** 17 / 17 * 2
