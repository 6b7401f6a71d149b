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
