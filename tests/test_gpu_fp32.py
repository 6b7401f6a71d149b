"""The FP32 form against the FP64 form on the same seeded state.

The FP32 form stores the prognostics in FP32 (s and qt as deviations from
FP64 level profiles), computes advection fluxes and FFTs in FP32 with a
compensated flux divergence for the scalars, and keeps the level means, the
tridiagonal recurrence and the saturation adjustment in FP64.  Bound: T and
qt within 1e-5 relative of FP64 after a few steps; velocities within 1e-3
of their scale.
"""
import numpy as np
import pytest
import torch

from tests.helpers import host_case

pytestmark = pytest.mark.gpu


def _run(nl, Gr, RS, Th, fields, precision, steps):
    import json
    from pycles_amd.Simulation3d import Simulation3d
    nl = json.loads(json.dumps(nl))
    nl["mi355x"]["precision"] = precision
    sim = Simulation3d().initialize(nl, Gr=Gr, Ref=RS, Th=Th, fields=fields)
    sim.run(max_steps=steps)
    torch.cuda.synchronize()
    return sim


@pytest.mark.parametrize("order", [(5, 5), (9, 5)])
def test_fp32_tracks_fp64(gpu, order):
    gw = (max(order) + 1) // 2
    nl, Gr, RS, Th, fields = host_case(32, 32, 48, gw=gw, order=order, seed=12)
    s64 = _run(nl, Gr, RS, Th, fields, "f64", 3)
    s32 = _run(nl, Gr, RS, Th, fields, "f32", 3)
    gw = Gr.dims.gw
    for n in ("s", "qt"):
        a = s32.PV.get_device_layout(n)[gw:-gw]
        b = s64.PV.get_device_layout(n)[gw:-gw]
        assert np.max(np.abs(a - b) / np.abs(b)) < 1e-5, n
    a = s32.DV.get_device_layout("temperature")[gw:-gw]
    b = s64.DV.get_device_layout("temperature")[gw:-gw]
    assert np.max(np.abs(a - b) / np.abs(b)) < 1e-5
    for n in ("u", "v", "w"):
        a = s32.PV.get_device_layout(n)
        b = s64.PV.get_device_layout(n)
        assert np.max(np.abs(a - b)) < 1e-3 * max(1.0, np.max(np.abs(b))), n


def test_fp32_thermo_clip_and_saturation(gpu):
    """One FP32 thermodynamics update on a state with a saturated share and a
    level of negative qt (clip_qt, thermodynamics_sa.h:213-220): T against the
    FP64 oracle within 1e-5, and the clipped level's stored qt at 1e-11 (its
    FP32 deviation from the level profile resolves ~1e-9)."""
    import json
    from pycles_amd.Simulation3d import Simulation3d
    from tests.helpers import oracle_sim
    from tests.test_gpu_parity import _saturating
    nl, Gr, RS, Th, fields = host_case(32, 24, 40, seed=4)
    fields = _saturating(Gr, fields)
    nl = json.loads(json.dumps(nl))
    nl["mi355x"]["precision"] = "f32"
    sim = Simulation3d().initialize(nl, Gr=Gr, Ref=RS, Th=Th, fields=fields)
    sim.Th.max_iter = 0
    sim.Th.update(sim.Gr, sim.Ref, sim.PV, sim.DV)
    torch.cuda.synchronize()
    osim = oracle_sim(Gr, RS, Th, fields, True)
    osim.thermo_update()
    gw = Gr.dims.gw
    assert osim.sa_stats.saturated > 0.05 * osim.npg
    T_o = Gr.to_device_layout(osim.dv["temperature"])[gw:-gw]
    T_g = sim.DV.get_device_layout("temperature")[gw:-gw]
    assert np.max(np.abs(T_g - T_o) / T_o) < 1e-5
    qt = sim.PV.get_device_layout("qt")
    assert np.all(np.abs(qt[gw + 2] - 1e-11) < 2e-9)
    others = np.delete(np.arange(gw, qt.shape[0] - gw), 2)
    qt_o = fields["qt"]
    assert np.max(np.abs(qt[others] - qt_o[others]) / np.abs(qt_o[others])) < 1e-6


def test_fp32_qt_clip_is_idempotent_across_chunks(gpu):
    """Clipped levels at the edge between two z chunks of the FP32
    thermodynamics kernel: a chunk reads its neighbour's edge level, which the
    owner may already have clipped and stored.  The stored deviation re-clips
    to exactly 1e-11 in every reader, so a second update on the clipped state
    gives the first update's outputs bit for bit (before, the stored FP32
    deviation could read back above 1e-11 and change T, b and N^2 there)."""
    import json
    from pycles_amd import _capi, _device as D
    from pycles_amd.Simulation3d import Simulation3d
    from tests.test_gpu_parity import _saturating
    nl, Gr, RS, Th, fields = host_case(32, 24, 40, seed=4)
    fields = _saturating(Gr, fields)
    nch, ln = _capi.launch_chunks(3, D.dims_of(Gr))
    assert nch > 1
    q = fields["qt"].copy()
    for k in (ln - 1, ln, ln + 1):
        q[k] = -1e-4
    fields["qt"] = q
    nl = json.loads(json.dumps(nl))
    nl["mi355x"]["precision"] = "f32"
    sim = Simulation3d().initialize(nl, Gr=Gr, Ref=RS, Th=Th, fields=fields)
    sim.Th.max_iter = 0
    out = []
    for _ in range(2):
        sim.Th.update(sim.Gr, sim.Ref, sim.PV, sim.DV, wt_overwrite=True)
        torch.cuda.synchronize()
        out.append((sim.DV.values.clone(), sim.PV.tend("w").clone(), sim.PV.field("qt").clone()))
    assert torch.equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1])
    assert torch.equal(out[0][2], out[1][2])
    qt = sim.PV.get_device_layout("qt")
    for k in (ln - 1, ln, ln + 1):
        assert np.all(qt[k] < 1e-11) and np.all(np.abs(qt[k] - 1e-11) < 2e-9)
