"""The benchmark configurations on their own workloads (sizes, spacings,
profiles, step counts and time steps of the five configurations):

  moist256   256x256x128 FP64, 10 steps at dt = 2 s, against the CPU oracle
             (u, v, w, s, qt, T, ql, p within 1e-12; the projected
             divergence at the oracle's level)
  moist512   512x512x256, FP32 with the learned first guess trained on 2^22
             pairs, against FP64 after 1 and 100 steps at dt = 1 s, with the
             noise floor of the comparison measured beside it
  moist1024  1024x1024x192, WENO9 scalars / WENO5 momentum, FP32 + learned
             guess, 20 steps at dt = 0.5 s, against FP64

The FP32 comparisons gate T and qt at 1e-5 relative.  Where the FP64 run
itself moves by more than that from a perturbation the FP32 form cannot avoid
(its own storage rounding of the initial state, run in FP64), the gate is
that floor: the FP32 form's error must stay within twice the error that the
rounding of its initial state alone causes."""
import numpy as np
import pytest
import torch

from tests.helpers import oracle_sim, rel_err

pytestmark = pytest.mark.gpu


def _state(name):
    import bench
    return bench.build_state(name)


def test_c2_moist256_fp64_against_the_oracle(gpu):
    from pycles_amd.Simulation3d import Simulation3d
    c, nl, Gr, RS, Th, fields = _state("moist256")
    assert (c["nx"], c["ny"], c["nz"], c["dt"], c["steps"]) == (256, 256, 128, 2.0, 10)
    sim = Simulation3d().initialize(nl, Gr=Gr, Ref=RS, Th=Th, fields=fields)
    sim.Th.max_iter = 0
    sim.run(max_steps=c["steps"])
    torch.cuda.synchronize()
    osim = oracle_sim(Gr, RS, Th, fields, True, dt=c["dt"])
    osim.begin_run()
    for _ in range(c["steps"]):
        osim.step()
    errs = {n: rel_err(sim.PV.get_device_layout(n), Gr.to_device_layout(osim.pv(n))) for n in osim.pv_names}
    for n in ("temperature", "ql", "dynamic_pressure"):
        errs[n] = rel_err(sim.DV.get_device_layout(n), Gr.to_device_layout(osim.dv[n]))
    print("moist256 FP64, 10 steps of 2 s, against the oracle:", errs, flush=True)
    assert sim.TS.t == osim.t == 20.0
    assert max(errs.values()) <= 1e-12, errs
    # the divergence left by the last projection, at the oracle's level
    gw = Gr.dims.gw
    sl = (slice(gw, -gw),) * 3
    div_o = np.abs(osim.divergence()[sl]).max()
    for n in ("u", "v", "w"):
        osim.pv(n)[...] = sim.PV.get_variable_array(n, Gr)
    div_g = np.abs(osim.divergence()[sl]).max()
    scale = max(np.abs(Gr.from_device_layout(fields[n])).max() for n in ("u", "v", "w")) / Gr.dims.dx[2]
    print(f"projected divergence: GPU {div_g:.3e}, oracle {div_o:.3e}, velocity/dz {scale:.3e}", flush=True)
    assert div_g <= max(10 * div_o, 1e-14 * scale)


def _fp32_gate(es, steps):
    e = es["errors_vs_fp64"]
    floor = es["noise_floor_fp64"]["fp64_from_fp32"]
    for key in ("after_1_steps", f"after_{steps}_steps"):
        for f in ("temperature", "qt"):
            err = e[key][f]["max_rel"]
            fl = floor[key][f]["max_rel"]
            print(f"{key} {f}: FP32 + learned {err:.3e}, FP64 from the FP32-rounded state {fl:.3e}", flush=True)
            assert err <= max(1e-5, 2.0 * fl), (key, f, err, fl)


def test_c3_c4_moist512_fp32_learned_against_fp64(gpu):
    import bench
    c, nl, Gr, RS, Th, fields = _state("moist512")
    assert (c["nx"], c["ny"], c["nz"], c["dt"], c["steps"]) == (512, 512, 256, 1.0, 100)
    guess, info = bench.train_guess(Th, 22, 40)
    assert info["pairs"] == 1 << 22
    print("learned guess:", {k: info[k] for k in ("heldout_rms_K", "heldout_max_K", "train_s")}, flush=True)
    assert info["heldout_rms_K"] < 0.05
    es = bench.error_study(nl, Gr, RS, Th, fields, guess, c["steps"], floors=True)
    _fp32_gate(es, c["steps"])
    # the per-step error is within the plain gate
    e1 = es["errors_vs_fp64"]["after_1_steps"]
    assert e1["temperature"]["max_rel"] <= 1e-5 and e1["qt"]["max_rel"] <= 1e-5


def test_c5_moist1024_weno9_fp32_learned_against_fp64(gpu):
    import bench
    c, nl, Gr, RS, Th, fields = _state("moist1024")
    assert (c["nx"], c["ny"], c["nz"], c["gw"], c["order"], c["dt"], c["steps"]) == \
        (1024, 1024, 192, 5, (9, 5), 0.5, 20)
    guess, _ = bench.train_guess(Th, 22, 40)
    es = bench.error_study(nl, Gr, RS, Th, fields, guess, c["steps"], floors=True)
    _fp32_gate(es, c["steps"])
