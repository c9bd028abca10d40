"""Pins of the FP64 surrogate-network oracle (oracle/mlp_oracle.py).

The surrogate has no counterpart in the reference, so the oracle is pinned by
mathematics: its analytic gradient against central finite differences of its
own loss, its BF16 rounding against hand-computed round-to-nearest-even
cases, and its forward pass against a closed form of a one-unit network.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_oracle as mo


def random_theta(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(mo.NUM_PARAMS) * scale).astype(np.float32)


def test_parameter_count_matches_the_header():
    assert mo.NUM_PARAMS == 33537
    header = (__import__("pathlib").Path(__file__).resolve().parents[1] / "include" / "bbrcit_hip.h").read_text()
    assert "#define BK_MLP_NUM_PARAMS 33537" in header


def test_bf16_round_to_nearest_even():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)          # BF16 spacing at 1
    cases = {
        1.0: 1.0,
        1.0 + 2.0 ** -8: 1.0,            # exact tie: round to even (1.0)
        1.0 + 3 * 2.0 ** -8: 1.0 + 2.0 ** -6,  # tie between 1+2^-7 (odd) and 1+2^-6 (even)
        1.0 + 2.0 ** -8 + 2.0 ** -20: 1.0 + 2.0 ** -7,  # above the tie
        -2.5: -2.5,
        3.0e38: 2.9993e38,               # near the top of the range stays finite
    }
    for x, want in cases.items():
        got = float(mo.bf16_round(np.float32(x)))
        if abs(want) < 1e38:
            assert got == pytest.approx(want, rel=0, abs=0), x
        else:
            assert np.isfinite(got) and abs(got - x) / x < 2 ** -8
    assert float(mo.bf16_round(one + ulp)) == float(one + ulp)
    # every BF16 value is a fixed point and has 16 zero low bits
    vals = mo.bf16_round(np.random.default_rng(0).standard_normal(1000).astype(np.float32))
    assert np.array_equal(mo.bf16_round(vals), vals)
    assert np.all((vals.view(np.uint32) & 0xFFFF) == 0)


def test_quantize_rounds_weights_not_biases():
    th = random_theta(1)
    q = mo.unpack(mo.quantize(th))
    p = mo.unpack(th)
    for k in ("W1", "W2", "W3", "W4"):
        assert np.array_equal(q[k], mo.bf16_round(p[k]).astype(np.float64))
    for k in ("b1", "b2", "b3", "b4"):
        assert np.array_equal(q[k], p[k].astype(np.float64))


def test_forward_closed_form_single_path():
    """Only unit 0 of each layer is live: y = w4 s(w3 s(w2 s(w1 . x + b1) + b2) + b3) + b4."""
    parts = {name: np.zeros(shape) for name, shape in mo.SIZES}
    parts["W1"][0] = [0.5, -1.25]
    parts["b1"][0] = 0.25
    parts["W2"][0, 0] = 1.5
    parts["b2"][0] = -0.5
    parts["W3"][0, 0] = -0.75
    parts["b3"][0] = 0.125
    parts["W4"][0] = 2.0
    parts["b4"][0] = 0.375
    wq = mo.pack(parts)
    x = np.array([[0.3, -0.6], [-1.0, 1.0]])
    s = lambda z: z / (1.0 + np.exp(-z))
    # units other than 0 see z = 0 -> s(0) = 0 and carry no weight forward
    h1 = s(0.5 * x[:, 0] - 1.25 * x[:, 1] + 0.25)
    h2 = s(1.5 * h1 - 0.5)
    h3 = s(-0.75 * h2 + 0.125)
    want = 2.0 * h3 + 0.375
    got, _ = mo.forward(wq, x)
    np.testing.assert_allclose(got, want, rtol=1e-15)


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    wq = mo.quantize(random_theta(2))
    x = rng.uniform(-1, 1, (16, 2))
    t = rng.standard_normal(16)
    loss, g = mo.loss_grad(wq, x, t)
    # every parameter of layer 1 and the output layer, a sample of the others
    idx = list(range(0, 384)) + list(range(33408, 33537))
    idx += list(rng.choice(np.arange(384, 33408), 300, replace=False))
    h = 1e-6
    for k in idx:
        e = np.zeros_like(wq)
        e[k] = h
        fd = (mo.loss_grad(wq + e, x, t)[0] - mo.loss_grad(wq - e, x, t)[0]) / (2 * h)
        assert fd == pytest.approx(g[k], rel=1e-5, abs=1e-9), k


def test_gradient_catches_a_planted_error():
    """A dropped SiLU' factor in the backward pass is visible to the check."""
    rng = np.random.default_rng(4)
    wq = mo.quantize(random_theta(5))
    x = rng.uniform(-1, 1, (8, 2))
    t = rng.standard_normal(8)
    _, g = mo.loss_grad(wq, x, t)
    saved = mo._dsilu
    try:
        mo._dsilu = lambda z: np.ones_like(z)
        _, g_bad = mo.loss_grad(wq, x, t)
    finally:
        mo._dsilu = saved
    k = 400  # a W2 entry
    h = 1e-6
    e = np.zeros_like(wq)
    e[k] = h
    fd = (mo.loss_grad(wq + e, x, t)[0] - mo.loss_grad(wq - e, x, t)[0]) / (2 * h)
    assert fd == pytest.approx(g[k], rel=1e-5, abs=1e-9)
    assert abs(g_bad[k] - fd) > 1e-3 * abs(fd)


def test_density_inverts_the_normalisation():
    wq = mo.quantize(random_theta(6, scale=0.05))
    q = np.array([[-4.0, 4.0], [0.0, 0.0], [1.5, -2.0]])
    mu, sigma, eps = -3.0, 2.0, 1e-12
    y, _ = mo.forward(wq, q / 4.0)
    np.testing.assert_allclose(mo.density(wq, q, -4.0, 4.0, mu, sigma, eps),
                               np.maximum(np.exp(mu + sigma * y) - eps, 0.0), rtol=1e-15)
