"""The closure oracle (oracle/closure_oracle.py) pinned on the CPU:
convolution against a brute-force loop, its two gradients as exact adjoints
(dot-product tests), GELU against known values and its derivative against
central differences, bf16 rounding against torch's conversion, and the
whole analytic gradient against finite differences."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.closure_oracle as co
from psyclonebench_amd.closure import init_params, layer_offsets


def brute_conv(x, w, b):
    ny, nx, cin = x.shape
    cout = w.shape[1]
    out = np.zeros((ny, nx, cout))
    for y in range(ny):
        for i in range(nx):
            for c in range(cout):
                s = b[c]
                for dy in range(5):
                    for dx in range(5):
                        yy, xx = y + dy - 2, i + dx - 2
                        if 0 <= yy < ny and 0 <= xx < nx:
                            s += np.dot(x[yy, xx, :], w[5 * dy + dx, c, :])
                out[y, i, c] = s
    return out


def test_conv_matches_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((4, 6, 3))
    w = rng.standard_normal((25, 2, 3))
    b = rng.standard_normal(2)
    np.testing.assert_allclose(co.conv(x, w, b), brute_conv(x, w, b), rtol=1e-12, atol=1e-12)


def test_conv_gradients_are_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((7, 9, 5))
    w = rng.standard_normal((25, 4, 5))
    d = rng.standard_normal((7, 9, 4))
    lhs = np.sum(co.conv(x, w, None) * d)
    assert np.isclose(lhs, np.sum(x * co.conv_input_grad(d, w)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(w * co.conv_weight_grad(d, x)), rtol=1e-12)


def test_gelu_values_and_derivative():
    import torch
    x = np.linspace(-9, 9, 721)
    # the tanh form as torch implements it, independently of the oracle
    ref = torch.nn.functional.gelu(torch.from_numpy(x), approximate="tanh").numpy()
    np.testing.assert_allclose(co.gelu(x), ref, rtol=1e-13, atol=1e-15)
    assert co.gelu(0.0) == 0.0
    # within 5e-4 of the erf form x Phi(x): Phi(1) = 0.8413447460685429
    assert abs(co.gelu(1.0) - 0.8413447460685429) < 5e-4
    assert abs(co.gelu(-1.0) + 0.15865525393145707) < 5e-4
    assert np.isclose(co.gelu(8.0), 8.0, rtol=1e-14)
    x = np.linspace(-5, 5, 101)
    h = 1e-6
    fd = (co.gelu(x + h) - co.gelu(x - h)) / (2 * h)
    np.testing.assert_allclose(co.gelu_grad(x), fd, rtol=1e-7, atol=1e-9)


def test_bf16_rounding():
    assert co.bf16(np.array([1.0]))[0] == 1.0
    assert co.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0               # tie -> even
    assert co.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    torch = pytest.importorskip("torch")
    x = np.random.default_rng(2).standard_normal(10000) * 10.0 ** np.random.default_rng(3).uniform(-6, 6, 10000)
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(co.bf16(x), ref)


def test_layout_matches_product():
    assert co.layer_offsets() == layer_offsets()


def test_analytic_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    p = init_params(5).astype(np.float64)
    # nonzero biases and output weights so every path carries gradient
    offs = co.layer_offsets()
    for l, (cin, cout) in enumerate(co.LAYERS):
        b0 = offs[l] + 25 * cin * cout
        p[b0:b0 + cout] = 0.1 * rng.standard_normal(cout)
    p[offs[3]:offs[4]] *= 10.0
    x = rng.standard_normal((6, 8, 4))
    t = 0.3 * rng.standard_normal((6, 8, 3))
    w = (1.0, 3.0, 0.5)   # channel weights: the weighted loss and its gradient
    _, g = co.loss_and_grad(p, x, t, quant=False, weights=w)
    picks = []
    for l, (cin, cout) in enumerate(co.LAYERS):
        ci_used = 4 if l == 0 else cin
        co_used = 3 if l == 3 else cout
        for _ in range(6):
            tap, c_o, c_i = rng.integers(25), rng.integers(co_used), rng.integers(ci_used)
            picks.append(offs[l] + (tap * cout + c_o) * cin + c_i)
        picks.append(offs[l] + 25 * cin * cout + rng.integers(co_used))
    h = 1e-5
    for k in picks:
        e = np.zeros_like(p)
        e[k] = h
        lp, _ = co.loss_and_grad(p + e, x, t, quant=False, weights=w)
        lm, _ = co.loss_and_grad(p - e, x, t, quant=False, weights=w)
        fd = (lp - lm) / (2 * h)
        assert np.isclose(g[k], fd, rtol=2e-5, atol=1e-10), (k, g[k], fd)


def test_padded_channels_get_no_gradient():
    rng = np.random.default_rng(6)
    p = init_params(7).astype(np.float64)
    _, g = co.loss_and_grad(p, rng.standard_normal((5, 8, 4)), rng.standard_normal((5, 8, 3)),
                            quant=True)
    w0 = g[:25 * 16 * 32].reshape(25, 32, 16)
    assert not np.any(w0[:, :, 4:])
    offs = co.layer_offsets()
    w3 = g[offs[3]:offs[3] + 25 * 32 * 16].reshape(25, 16, 32)
    assert not np.any(w3[:, 3:, :])
    assert not np.any(g[offs[3] + 25 * 32 * 16 + 3:offs[4]])
