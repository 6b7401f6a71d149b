"""The learned first guess of the saturation adjustment on the GPU: training
pairs against the oracle's secant, the library's gradient, forward and packed
(bf16 matrix-core) inference against the CPU oracle of the network, the first
Adam step, training progress, and the thermodynamics with the guess against
the FP64 oracle, with the iterations it saves over the plain secant.
"""
import json

import numpy as np
import pytest
import torch

from oracle import mlp as M
from oracle import oracle as O
from tests.helpers import host_case, oracle_sim

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(gpu):
    nl, Gr, RS, Th, fields = host_case(32, 32, 48, seed=21)
    return nl, Gr, RS, Th, fields


@pytest.fixture(scope="module")
def trained(case):
    from pycles_amd.LearnedGuess import SaturationGuess
    Th = case[3]
    g = SaturationGuess(Th.CC, n_pairs=1 << 20, batch=4096, epochs=24, seed=7).fit()
    torch.cuda.synchronize()
    return g


def _host(t):
    return t.detach().cpu().numpy()


def test_training_pairs_are_secant_solutions(case):
    from pycles_amd.LearnedGuess import SaturationGuess
    Th = case[3]
    g = SaturationGuess(Th.CC, n_pairs=1 << 12, seed=3)
    d = _host(g.pairs(4096, 3, saturated_only=False))
    assert np.all(np.isfinite(d))
    assert d[:, 1].min() >= 0 and d[:, 1].max() <= 0.025
    assert d[:, 2].min() >= 5e4 and d[:, 2].max() <= 1.02e5
    lt = O.Lookup(Th.CC.T, Th.CC.pv)
    sat = 0
    for row in d[:400]:
        T, qv, ql, qi, it = O.eos_sa(lt, row[2], row[0], row[1])
        sat += it > 0
        assert abs(T - row[3]) <= 1e-3, (row, T)  # a last-bit flip of the 1e-3 K stop at most
    assert 50 < sat < 390  # both branches are sampled
    ds = _host(g.pairs(512, 4, saturated_only=True))
    for row in ds[:200]:
        T, qv, ql, qi, it = O.eos_sa(lt, row[2], row[0], row[1])
        assert it > 0 and abs(T - row[3]) <= 1e-3


def test_gradient_matches_oracle(case):
    # at the initial weights, where the per-sample contributions do not cancel
    from pycles_amd.LearnedGuess import SaturationGuess
    g = SaturationGuess(case[3].CC, n_pairs=1 << 14, seed=11).generate()
    data = g.data[:1024]
    grad, loss = g.gradient(data)
    p = _host(g.params).astype(np.float64)
    x, t = M.normalise(_host(data), g.normalisation())
    loss_o, grad_o = M.loss_and_grad(p, x, t)
    assert abs(float(loss.item()) - loss_o) <= 1e-4 * loss_o + 1e-9
    gg = _host(grad).astype(np.float64)
    for a, b in ((M.OW1, M.OB1), (M.OB1, M.OW2), (M.OW2, M.OB2), (M.OB2, M.OW3), (M.OW3, M.OB3)):
        err = np.abs(gg[a:b] - grad_o[a:b]).max() / np.abs(grad_o[a:b]).max()
        assert err < 1e-4, (a, err)
    assert abs(gg[M.OB3] - grad_o[M.OB3]) <= 1e-4 * np.abs(grad_o).max()


def test_forward_and_packed_inference_match_oracle(trained):
    g = trained
    held = g.pairs(8192, 99)
    p = _host(g.params).astype(np.float64)
    nm = g.normalisation()
    x, _ = M.normalise(_host(held), nm)
    y32 = _host(g.predict(held, packed=False)).astype(np.float64)
    yo = nm["mu_T"] + nm["sd_T"] * M.forward(p, x)[0]
    assert np.abs(y32 - yo).max() < 2e-4  # K, FP32 arithmetic
    ybf = _host(g.predict(held, packed=True)).astype(np.float64)
    yob = nm["mu_T"] + nm["sd_T"] * M.infer_bf16(p, x)
    T = _host(held)[:, 3]
    err32 = y32 - T
    errbf = ybf - T
    d = ybf - yob
    print(json.dumps({"guess_rms_K_fp32": float(np.sqrt(np.mean(err32 ** 2))),
                      "guess_max_K_fp32": float(np.abs(err32).max()),
                      "guess_rms_K_bf16": float(np.sqrt(np.mean(errbf ** 2))),
                      "guess_max_K_bf16": float(np.abs(errbf).max()),
                      "mfma_vs_oracle_rms_K": float(np.sqrt(np.mean(d ** 2))),
                      "mfma_vs_oracle_max_K": float(np.abs(d).max())}))
    # The hidden activations are rounded to bf16 on both sides; the device's
    # fast tanh moves a few of them across a bf16 rounding boundary (one step
    # is 2^-8 relative), which shifts T by up to a few 1e-2 K (measured:
    # 9e-4 K RMS, 0.03 K at most).  A wrong lane layout or K order of the
    # matrix-core operands would give errors of kelvins.
    assert np.sqrt(np.mean(d ** 2)) < 2.5e-3
    assert np.abs(d).max() < 0.06


def test_first_adam_step(case):
    from pycles_amd.LearnedGuess import SaturationGuess
    Th = case[3]
    g = SaturationGuess(Th.CC, n_pairs=4096, batch=4096, epochs=1, lr_max=1e-3, lr_min=1e-3, seed=5)
    g.generate()
    p0 = _host(g.params).copy()
    grad, _ = g.gradient(g.data)
    g.train(epochs=1)
    dp = _host(g.params) - p0
    gg = _host(grad)
    big = np.abs(gg) > 1e-4 * np.abs(gg).max()
    # first Adam step: m_hat = g, v_hat = g^2, so the step is lr * sign(g)
    assert np.allclose(dp[big], -1e-3 * np.sign(gg[big]), rtol=1e-3, atol=0)


def test_training_reduces_the_loss(trained):
    h = trained.loss_history
    assert len(h) == 24 and all(np.isfinite(h))
    assert h[-1] < 0.02 * h[0], h
    held = trained.pairs(1 << 14, 1234)
    T = _host(held)[:, 3]
    err = _host(trained.predict(held)).astype(np.float64) - T
    assert np.sqrt(np.mean(err ** 2)) < 0.5, np.sqrt(np.mean(err ** 2))


def _saturating(Gr, fields, seed=3):
    d = Gr.dims
    rng = np.random.default_rng(seed)
    f = dict(fields)
    s = fields["s"].copy()
    s[d.gw:d.gw + d.n[2]] -= rng.uniform(0.0, 25.0, size=(d.n[2], d.n[1], d.n[0]))
    f["s"] = s
    return f


def _thermo(nl, Gr, RS, Th, fields, mlp, newton):
    """One FP32 thermodynamics update; returns T, ql and a copy of the counters."""
    from pycles_amd.Simulation3d import Simulation3d
    nl = json.loads(json.dumps(nl))
    nl["mi355x"]["precision"] = "f32"
    sim = Simulation3d().initialize(nl, Gr=Gr, Ref=RS, Th=Th, fields=fields)
    sim.Th.mlp = mlp
    sim.Th.newton_steps = newton
    sim.Th.max_iter = 0
    sim.Th._prepare(sim.Gr)
    sim.Th.counters.zero_()
    sim.Th.update(sim.Gr, sim.Ref, sim.PV, sim.DV)
    torch.cuda.synchronize()
    out = (sim.DV.get_device_layout("temperature"), sim.DV.get_device_layout("ql"),
           sim.Th.counters.cpu().numpy().copy())
    sim.Th.mlp = None
    return out


@pytest.mark.parametrize("newton", [1, 2])
def test_thermo_with_guess_matches_fp64_and_saves_iterations(case, trained, newton):
    nl, Gr, RS, Th, fields = case
    fields = _saturating(Gr, fields)
    plain = _thermo(nl, Gr, RS, Th, fields, None, 1)
    learned = _thermo(nl, Gr, RS, Th, fields, trained, newton)
    osim = oracle_sim(Gr, RS, Th, fields, True)
    osim.thermo_update()
    gw = Gr.dims.gw
    T_o = Gr.to_device_layout(osim.dv["temperature"])[gw:-gw]
    ql_o = Gr.to_device_layout(osim.dv["ql"])[gw:-gw]
    for T, ql, _ in (plain, learned):
        assert np.max(np.abs(T[gw:-gw] - T_o) / T_o) < 1e-5
        # the secant reports qv at its last-but-one iterate (within 1e-3 K);
        # the Newton path reports qv*(T) at its final T
        assert np.max(np.abs(ql[gw:-gw] - ql_o)) < 2e-6
    cp, cl = plain[2], learned[2]
    sat = int(cp[0])
    assert sat > 0.05 * T_o.size and int(cl[0]) == sat
    secant_per_cell = cp[1] / sat
    newton_per_cell = cl[1] / sat
    fallback = cl[3] / sat
    print(json.dumps({"newton": newton, "saturated": sat, "secant_iters_per_cell": secant_per_cell,
                      "newton_iters_per_cell": newton_per_cell, "fallback_fraction": fallback,
                      "plain_fallbacks": int(cp[3])}))
    # the FP32 secant's results pass the FP64 residual check: a check that
    # rejected good results would send every cell to the FP64 secant, with
    # the same values and only the time lost
    assert cp[3] <= 0.01 * sat
    # a cell whose guess fails the check also runs the FP32 secant
    assert newton <= newton_per_cell <= newton + fallback * (secant_per_cell + 5)
    assert secant_per_cell > newton_per_cell
    # one correction squares the guess's error (a few 1e-2 K typical, tenths
    # of a kelvin at worst); the FP64 residual test sends the rare miss back
    assert fallback < (0.05 if newton == 1 else 0.005)
