"""Pins the CPU oracle of the warm-start network: bf16 rounding on known
values, the backward pass against central finite differences of the
forward pass (rounding off, float64), and Adam against its closed form."""
import numpy as np
import pytest

from oracle import mlp


def test_bf16_round_known_values():
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 0.0, 3.0e38, 1.0 + 2.0 ** -9],
                    dtype=np.float64)
    # ties go to even: 1+2^-8 -> 1, 1+3*2^-8 -> 1+2^-6; 1+2^-7 is exact
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, 0.0, np.float32(3.0e38), 1.0])
    got = mlp.bf16_round(vals)
    want[6] = 2.999393627791264e38  # nearest bf16 to float32(3e38)
    np.testing.assert_array_equal(got[[0, 1, 2, 3, 4, 5, 7]], want[[0, 1, 2, 3, 4, 5, 7]])
    assert abs(got[6] - 3.0e38) / 3.0e38 < 2.0 ** -8
    # every result has at most 8 significant bits
    r = mlp.bf16_round(np.random.default_rng(1).normal(size=1000))
    assert np.all(r.astype(np.float32).view(np.uint32) & 0xFFFF == 0)
    assert np.isnan(mlp.bf16_round(np.array([np.nan]))[0])


def test_gelu_known_values():
    assert mlp.gelu(np.array([0.0]))[0] == 0.0
    # gelu(z) = z * Phi(z); Phi(1) = 0.8413447460685429, Phi(3) = 0.9986501019683699 (normal tables)
    np.testing.assert_allclose(mlp.gelu(np.array([1.0, -1.0, 3.0])),
                               [0.8413447460685429, -(1 - 0.8413447460685429), 3 * 0.9986501019683699], rtol=1e-14)
    z = np.linspace(-4, 4, 17)
    h = 1e-6
    fd = (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(mlp.gelu_grad(z), fd, atol=1e-8)


@pytest.mark.parametrize("n_in,hid,n_out,rows", [(3, 16, 5, 7), (2, 24, 3, 4)])
def test_backward_matches_finite_differences(n_in, hid, n_out, rows):
    rng = np.random.default_rng(7)
    p = rng.normal(scale=0.4, size=mlp.num_params(n_in, hid, n_out))
    x = rng.uniform(-1, 1, size=(rows, n_in))
    t = rng.normal(size=(rows, n_out))
    loss, g = mlp.loss_and_grad(p, x, t, n_in, hid, n_out, bf16=False)
    h = 1e-6
    idx = rng.choice(p.size, size=60, replace=False)
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = mlp.loss_and_grad(p + e, x, t, n_in, hid, n_out, bf16=False)
        lm, _ = mlp.loss_and_grad(p - e, x, t, n_in, hid, n_out, bf16=False)
        assert abs((lp - lm) / (2 * h) - g[i]) < 1e-7 * max(1.0, abs(g[i])), i
    # a planted mistake (dropping the GELU derivative of layer 2) is caught
    W1, b1, W2, b2, W3, b3 = mlp.unpack(p, n_in, hid, n_out)
    y, (xx, z1, a1, z2, a2) = mlp.forward(p, x, n_in, hid, n_out, bf16=False, keep=True)
    d3 = 2 * (y - t) / y.size
    d2_wrong = d3 @ W3
    assert np.max(np.abs(d2_wrong.T @ a1 - g[hid * n_in + hid:hid * n_in + hid + hid * hid].reshape(hid, hid))) > 1e-6


def test_bf16_forward_close_to_f64():
    rng = np.random.default_rng(3)
    p = rng.normal(scale=0.2, size=mlp.num_params(3, 32, 8))
    x = rng.uniform(-1, 1, size=(10, 3))
    y64 = mlp.forward(p, x, 3, 32, 8, bf16=False)
    y16 = mlp.forward(p, x, 3, 32, 8, bf16=True)
    assert 0 < np.max(np.abs(y64 - y16)) < 0.05


def test_adam_step_closed_form():
    # first step moves every coordinate by lr * g / (|g| + eps)
    g = np.array([0.5, -2.0, 1e-3])
    p, m, v = mlp.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.1)
    np.testing.assert_allclose(p, -0.1 * g / (np.abs(g) + 1e-8), rtol=1e-12)
