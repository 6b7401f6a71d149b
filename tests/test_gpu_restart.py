"""Checkpoint and restart on the GPU: a run that writes a restart file after
step 2 and goes on to step 4, against a second run started from that file
and taken from step 2 to step 4.  The two end states must be identical bit
for bit (values, diagnostics, time and time step), in FP64, in FP32 and in
FP32 with the learned first guess (Restart.pyx:92-152,
PrognosticVariables.pyx:269-357)."""
import copy
import os

import numpy as np
import pytest
import torch

from tests.helpers import host_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,learned", [("f64", False), ("f32", False), ("f32", True)])
def test_resume_is_bitwise(gpu, tmp_path, precision, learned):
    from pycles_amd.ReferenceState import ReferenceState
    from pycles_amd.Simulation3d import Simulation3d
    from pycles_amd.ThermodynamicsSA import ThermodynamicsFactory
    nl, Gr, RS, Th, fields = host_case(32, 16, 24, seed=31)
    nl = copy.deepcopy(nl)
    nl["mi355x"]["precision"] = precision
    nl["meta"] = {"uuid": "restart-test-12345", "simname": "resume"}
    nl["output"] = {"output_root": str(tmp_path) + "/"}
    nl["restart"] = {"frequency": 2.0}
    guess = None
    if learned:
        from pycles_amd.LearnedGuess import SaturationGuess
        guess = SaturationGuess(Th.CC, n_pairs=1 << 14, epochs=3).fit()
    sim = Simulation3d().initialize(nl, Gr=Gr, Ref=RS, Th=Th, fields=fields)
    sim.Th.mlp = guess
    sim.run(max_steps=4)
    torch.cuda.synchronize()
    path = os.path.join(sim.Restart.restart_path, "2")
    assert os.path.exists(os.path.join(path, "0.pcr")), os.listdir(sim.Restart.restart_path)
    want = (sim.PV.values.clone(), sim.DV.values.clone(), sim.TS.t, sim.TS.dt)
    assert sim.TS.t == 4.0

    nl2 = copy.deepcopy(nl)
    nl2["restart"] = {"frequency": 2.0, "init_from": True, "input_path": path}
    nl2["meta"]["uuid"] = "restart-test-54321"
    Th2 = ThermodynamicsFactory(nl2)
    RS2 = ReferenceState(Gr)
    sim2 = Simulation3d().initialize(nl2, Gr=Gr, Ref=RS2, Th=Th2)
    assert sim2.TS.t == 2.0
    assert (sim2.Th.mlp is not None) == learned
    assert np.array_equal(RS2.p0_half, RS.p0_half) and np.array_equal(RS2.alpha0, RS.alpha0)
    sim2.run(max_steps=2)
    torch.cuda.synchronize()
    assert sim2.TS.t == want[2] and sim2.TS.dt == want[3]
    assert torch.equal(sim2.PV.values, want[0])
    assert torch.equal(sim2.DV.values, want[1])
    for n in ("s", "qt"):
        if precision == "f32":
            assert torch.equal(sim2.PV.base[n], sim.PV.base[n])
