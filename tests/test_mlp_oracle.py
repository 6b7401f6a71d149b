"""Pins of the CPU oracle of the learned saturation guess: its analytic
gradient against central finite differences, and its bf16 rounding against
the format's definition (8 significant bits, nearest, ties to even)."""
import numpy as np

from oracle import mlp as M


def _params(rng):
    p = np.zeros(M.NPARAM)
    p[M.OW1:M.OB1] = rng.uniform(-0.3, 0.3, M.NH * M.NIN)
    p[M.OB1:M.OW2] = rng.uniform(-0.1, 0.1, M.NH)
    p[M.OW2:M.OB2] = rng.uniform(-0.2, 0.2, M.NH * M.NH)
    p[M.OB2:M.OW3] = rng.uniform(-0.1, 0.1, M.NH)
    p[M.OW3:M.OB3] = rng.uniform(-0.3, 0.3, M.NH)
    p[M.OB3] = 0.05
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng)
    x = rng.normal(size=(37, 3))
    t = rng.normal(size=37)
    _, g = M.loss_and_grad(p, x, t)
    h = 1e-6
    idx = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False) for a, b in
                          ((M.OW1, M.OB1), (M.OB1, M.OW2), (M.OW2, M.OB2), (M.OB2, M.OW3), (M.OW3, M.OB3))]
                         + [np.array([M.OB3])])
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (M.loss_and_grad(pp, x, t)[0] - M.loss_and_grad(pm, x, t)[0]) / (2 * h) / len(t)
        assert abs(fd - g[q]) <= 1e-7 * max(1.0, abs(g[q])), (q, fd, g[q])


def test_forward_shapes_and_loss():
    rng = np.random.default_rng(1)
    p = _params(rng)
    x = rng.normal(size=(5, 3))
    y, _ = M.forward(p, x)
    loss, _ = M.loss_and_grad(p, x, y)
    assert y.shape == (5,) and loss == 0.0


def test_bf16_rounding():
    two = np.float64(2.0)
    vals = np.array([1.0, 1.0 + two ** -8, 1.0 + 3 * two ** -9, 1.0 + two ** -7 + two ** -8, -3.0, 0.0,
                     65504.0, 1e-30], dtype=np.float32)
    r = M.bf16_round(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                        # tie -> even (mantissa 0)
    assert r[2] == 1.0 + two ** -7            # above the tie -> up
    assert r[3] == 1.0 + two ** -6            # tie with odd mantissa -> up to even
    assert r[4] == -3.0 and r[5] == 0.0
    for v, q in zip(vals, r):
        if v != 0:
            assert abs(q - v) <= abs(v) * two ** -8
            m = np.frexp(q)[0] * 2 ** 8           # at most 8 significant bits
            assert m == np.round(m)


def test_bf16_inference_is_the_forward_pass_to_bf16_precision():
    """infer_bf16 (what the packed matrix-core inference computes) differs from
    the float64 forward pass only by the bf16 roundings of W1, TANH_C W2, W3
    and the first hidden activations: each is 2^-9 relative at most, so the
    output moves by a few 1e-3 of its scale and not more; and TANH_C is the
    FP32 value of 2 log2(e), the factor of tanh(x) = 1 - 2/(2^(TANH_C x) + 1)."""
    rng = np.random.default_rng(4)
    p = _params(rng)
    x = rng.normal(size=(4096, 3))
    y64 = M.forward(p, x)[0]
    ybf = M.infer_bf16(p, x)
    scale = np.abs(y64).max()
    err = np.abs(ybf - y64)
    assert err.max() < 1e-2 * scale
    assert np.sqrt(np.mean(err ** 2)) < 3e-3 * scale
    assert err.max() > 0.0  # the roundings are there
    assert abs(M.TANH_C - 2.0 / np.log(2.0)) < 1e-7 * M.TANH_C
    u = np.linspace(-3.0, 3.0, 13)
    assert np.allclose(1.0 - 2.0 / (np.exp2(M.TANH_C * u) + 1.0), np.tanh(u), rtol=0, atol=1e-6)
