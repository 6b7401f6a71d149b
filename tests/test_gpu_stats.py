"""On-device statistics against the oracle's restatement on the same state:
per-level moments, extremes, products and conditional means
(ParallelMPI.pyx:196-540) in FP64 and FP32 (with the scalars' level
profiles), the fields of ThermodynamicsSA.stats_io and the liquid-water
statistics (ThermodynamicsSA.pyx:256-492), and the statistics output of a
run with a stats_io namelist section.

Tolerances: the device sums in a fixed tree order, the reference serially,
so means agree to a few ulps of the sum (1e-13 relative to the level's
scale); extremes are exact; the cloud fraction adds 1/n per cloudy cell in
the reference and multiplies the count by 1/n here (a few ulps)."""
import copy

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.helpers import host_case, oracle_sim

pytestmark = pytest.mark.gpu


def _sim(nl, Gr, RS, Th, fields, precision):
    from pycles_amd.Simulation3d import Simulation3d
    nl = copy.deepcopy(nl)
    nl["mi355x"]["precision"] = precision
    return Simulation3d().initialize(nl, Gr=Gr, Ref=RS, Th=Th, fields=fields)


def _close(a, b, rtol=1e-13):
    scale = np.maximum(np.abs(b), 1e-300)
    return np.all(np.abs(a - b) <= rtol * np.maximum(scale, np.abs(b).max(axis=-1, keepdims=True) * 1e-3))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_horizontal_stats(gpu, precision):
    from tests.test_gpu_parity import _saturating
    nl, Gr, RS, Th, fields = host_case(24, 16, 20, seed=41)
    fields = _saturating(Gr, fields)
    sim = _sim(nl, Gr, RS, Th, fields, precision)
    sim.begin_run()
    torch.cuda.synchronize()
    d = O.make_dims(*Gr.dims.n, Gr.dims.gw, *Gr.dims.dx)
    PV, Pa = sim.PV, sim.Pa
    tol = 1e-13
    for name in ("u", "w", "s", "qt"):
        st = Pa.horizontal_stats(Gr, PV.field(name), PV.precision, PV.base.get(name))
        arr = PV.get_variable_array(name, Gr)
        ref = O.horizontal_stats(d, arr)
        # a sum of n terms in two orders differs by at most ~n eps times the
        # sum of their magnitudes: the level means of |v|^p set the scale
        # (the means of w are cancellation noise, ~1e-18 m/s)
        mag = O.horizontal_stats(d, np.abs(arr))
        for r in range(3):
            assert np.all(np.abs(st[r] - ref[r]) <= tol * mag[r]), (name, r, np.abs(st[r] - ref[r]).max())
        assert np.array_equal(st[3], ref[3]) and np.array_equal(st[4], ref[4]), name
    # qt*s product and a conditional mean over the saturated cells
    st = Pa.horizontal_stats(Gr, PV.field("qt"), PV.precision, PV.base.get("qt"), PV.field("s"), PV.base.get("s"))
    ref = O.horizontal_stats(d, PV.get_variable_array("qt", Gr), PV.get_variable_array("s", Gr))
    assert np.all(np.abs(st[0] - ref[0]) <= tol * np.abs(ref[0]))
    ql = sim.DV.field("ql")
    mask = (ql > 0).to(ql.dtype)
    st = Pa.horizontal_stats(Gr, sim.DV.field("temperature"), sim.DV.precision, mask=mask)
    ref = O.horizontal_stats(d, sim.DV.get_variable_array("temperature", Gr),
                             mask=Gr.from_device_layout(mask.to(torch.float64).cpu().numpy()))
    assert np.array_equal(st[5], ref[5]) and ref[5].sum() > 0
    assert np.all(np.abs(st[0] - ref[0]) <= tol * np.maximum(np.abs(ref[0]), 1.0))
    assert np.array_equal(st[3], ref[3]) and np.array_equal(st[4], ref[4])


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_sa_stat_fields_and_liquid_stats(gpu, precision):
    from tests.test_gpu_parity import _saturating
    nl, Gr, RS, Th, fields = host_case(32, 16, 24, seed=42)
    fields = _saturating(Gr, fields)
    sim = _sim(nl, Gr, RS, Th, fields, precision)
    sim.begin_run()
    NS = sim.StatsIO
    NS.open_files()
    sim.Th.stats_io(sim.Gr, sim.Ref, sim.PV, sim.DV, NS, sim.Pa)
    cur = NS.close_files()
    d = O.make_dims(*Gr.dims.n, Gr.dims.gw, *Gr.dims.dx)
    g = lambda n, src: src.get_variable_array(n, Gr)  # noqa: E731
    lt = O.Lookup(Th.CC.T, Th.CC.pv)
    thetas, theta, rh = O.sa_stat_fields(RS.p0_half, lt, g("s", sim.PV), g("qt", sim.PV),
                                         g("temperature", sim.DV), g("qv", sim.DV))
    gw = Gr.dims.gw
    # FP32 stores the derived fields in FP32 before they are averaged
    tol = 1e-13 if precision == "f64" else 1e-6
    for name, fld in (("thetas", thetas), ("theta", theta), ("rh", rh)):
        ref = O.horizontal_stats(d, fld)
        assert np.all(np.abs(cur["profile/" + name + "_mean"] - ref[0, gw:-gw]) <= tol * np.abs(ref[0, gw:-gw])), name
        assert np.all(np.abs(cur["profile/" + name + "_max"] - ref[3, gw:-gw]) <= tol * np.abs(ref[3, gw:-gw])), name
    cf, cover, cb, ct, lwp = O.liquid_stats(d, RS.rho0_half, g("ql", sim.DV))
    assert cover > 0.05
    assert np.allclose(cur["profile/cloud_fraction"], cf, rtol=1e-13, atol=0)
    assert float(cur["ts/cloud_fraction"]) == cover
    assert float(cur["ts/cloud_base"]) == cb and float(cur["ts/cloud_top"]) == ct
    assert float(cur["ts/lwp"]) == pytest.approx(lwp, rel=1e-13)


def test_run_writes_statistics(gpu, tmp_path):
    nl, Gr, RS, Th, fields = host_case(16, 16, 16, seed=43)
    nl = copy.deepcopy(nl)
    nl["meta"] = {"uuid": "stats-run-00001", "simname": "stats"}
    nl["output"] = {"output_root": str(tmp_path) + "/"}
    nl["stats_io"] = {"frequency": 2.0, "stats_dir": "stats"}
    from pycles_amd.Simulation3d import Simulation3d
    sim = Simulation3d().initialize(nl, Gr=Gr, Ref=RS, Th=Th, fields=fields)
    sim.run(max_steps=4)
    assert len(sim.StatsIO.written) == 2
    with np.load(sim.StatsIO.written[-1], allow_pickle=False) as z:
        keys = set(z.files)
        assert float(z["t"]) == 4.0
        for k in ("profile/u_mean", "profile/qt_mean3", "profile/temperature_max", "profile/thetas_mean",
                  "profile/rh_min", "profile/cloud_fraction", "ts/lwp", "profile/qt_s_product_mean"):
            assert k in keys, k
        assert z["profile/u_mean"].shape == (16,)
