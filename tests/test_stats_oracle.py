"""Pins of the oracle's statistics restatement (ParallelMPI.pyx:196-540,
ThermodynamicsSA.pyx:256-492) by closed forms: moments of constant and
two-valued levels, the conditional mean's divisor, the extreme values'
initial bounds, thetas at the reference entropy, theta at the reference
pressure, a saturated relative humidity of one, and the cloud statistics of
a constructed cloud layer."""
import numpy as np
import pytest

from oracle import oracle as O
from pycles_amd import parameters as P


def _dims():
    return O.make_dims(6, 5, 4, 3, 50.0, 50.0, 20.0)


def test_moments_of_simple_levels():
    d = _dims()
    n0, n1, n2 = d.nlg
    a = np.zeros((n0, n1, n2))
    gw = d.gw
    a[:, :, 0] = 2.5                                    # constant level
    a[gw:gw + 3, :, 1] = 1.0                            # half the interior 1, half -1
    a[gw + 3:, :, 1] = -1.0
    st = O.horizontal_stats(d, a)
    assert st[0, 0] == pytest.approx(2.5, rel=1e-15) and st[1, 0] == pytest.approx(6.25, rel=1e-15)
    assert st[2, 0] == pytest.approx(15.625, rel=1e-15) and st[3, 0] == 2.5 and st[4, 0] == 2.5
    assert abs(st[0, 1]) < 1e-15 and st[1, 1] == pytest.approx(1.0) and abs(st[2, 1]) < 1e-15
    assert st[3, 1] == 1.0 and st[4, 1] == -1.0 and st[5, 1] == 30.0
    # the extremes start from -9e12 / 9e12 (ParallelMPI.pyx:345, 377)
    big = np.full_like(a, -1e13)
    assert O.horizontal_stats(d, big)[3, 2] == -9e12
    # product of two fields and a conditional mean divides by the mask count
    b = np.full_like(a, 3.0)
    assert O.horizontal_stats(d, a, b)[0, 0] == pytest.approx(7.5, rel=1e-15)
    m = np.zeros_like(a)
    m[gw:gw + 3, :, 1] = 1.0
    st = O.horizontal_stats(d, a, mask=m)
    assert st[0, 1] == pytest.approx(1.0, rel=1e-15) and st[5, 1] == 15.0
    assert st[0, 3] == 0.0  # empty mask: divisor max(0, 1)


def test_sa_stat_fields_closed_forms():
    from pycles_amd.Thermodynamics import ClausiusClapeyron, LatentHeat
    cc = ClausiusClapeyron().initialize(None, LatentHeat())
    lt = O.Lookup(cc.T, cc.pv)
    shape = (2, 2, 3)
    s = np.full(shape, P.sd_tilde)
    qt = np.zeros(shape)
    T = np.full(shape, 290.0)
    qv = np.zeros(shape)
    p0 = np.array([P.p_tilde, 9e4, 8e4])
    thetas, theta, rh = O.sa_stat_fields(p0, lt, s, qt, T, qv)
    assert np.allclose(thetas, P.T_tilde, rtol=1e-15)
    assert np.allclose(theta[..., 0], 290.0, rtol=1e-15)
    assert np.allclose(theta[..., 1], 290.0 * (P.p_tilde / 9e4) ** P.kappa, rtol=1e-14)
    # vapour at saturation: qv = qv*(T) gives rh = 1
    pvs = lt(290.0)
    qt2 = np.full(shape, 0.02)
    qvs = P.eps_v * (1 - 0.02) * pvs / (p0 - pvs)
    _, _, rh = O.sa_stat_fields(p0, lt, s, qt2, T, np.broadcast_to(qvs, shape).copy())
    assert np.allclose(rh, 1.0, rtol=1e-13)


def test_liquid_stats_of_a_cloud_layer():
    d = O.make_dims(4, 5, 6, 3, 50.0, 50.0, 20.0)
    gw = d.gw
    ql = np.zeros(tuple(d.nlg))
    # interior levels 2 and 3 cloudy in 10 of the 20 columns, ql = 1e-3
    ql[gw:gw + 2, gw:gw + 5, gw + 2:gw + 4] = 1e-3
    rho0_half = np.linspace(1.2, 0.9, d.nlg[2])
    cf, cover, cb, ct, lwp = O.liquid_stats(d, rho0_half, ql)
    assert np.allclose(cf, [0, 0, 0.5, 0.5, 0, 0], rtol=1e-14, atol=0)
    assert cover == 0.5
    assert cb == pytest.approx(2.5 * 20.0) and ct == pytest.approx(3.5 * 20.0)
    # interior levels 2, 3 weighted with the ghosted profile's entries 2, 3
    assert lwp == pytest.approx(0.5 * (rho0_half[2] + rho0_half[3]) * 1e-3 * 20.0, rel=1e-14)
    cf, cover, cb, ct, lwp = O.liquid_stats(d, rho0_half, np.zeros_like(ql))
    assert cover == 0.0 and cb == 99999.9 and ct == -99999.9 and lwp == 0.0
