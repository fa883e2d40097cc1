"""GPU warm-start network kernels (bf16 MFMA forward/backward, Adam) against
the float64 CPU oracle with the same bf16 operand rounding.

Tolerances: the GPU accumulates in f32 and may round an activation to the
neighbouring bf16 value where the f32 and f64 activations straddle a
rounding boundary, so outputs are compared to 1e-3 of their scale and
gradients by relative norm per parameter block (2e-3)."""
import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def _net(hip, n_in, hid, n_out, seed=1):
    from logosq_amd.warmstart import WarmStartNet
    return WarmStartNet(n_in, hid, n_out, seed=seed)


def _blocks(n_in, hid, n_out):
    sizes = [hid * n_in, hid, hid * hid, hid, n_out * hid, n_out]
    names = ["W1", "b1", "W2", "b2", "W3", "b3"]
    o = 0
    for nm, s in zip(names, sizes):
        yield nm, slice(o, o + s)
        o += s


@pytest.mark.parametrize("n_in,hid,n_out,rows", [(3, 256, 320, 64), (3, 48, 5, 37), (2, 16, 1, 1)])
def test_forward_matches_oracle(hip, n_in, hid, n_out, rows):
    print("mlp forward", n_in, hid, n_out, rows, flush=True)
    net = _net(hip, n_in, hid, n_out)
    p = net.get_params().astype(np.float64)
    assert p.size == omlp.num_params(n_in, hid, n_out)
    x = np.random.default_rng(5).uniform(-1, 1, size=(rows, n_in)).astype(np.float32)
    y = net.forward(x)
    want = omlp.forward(p, x.astype(np.float64), n_in, hid, n_out, bf16=True)
    scale = max(1.0, np.max(np.abs(want)))
    assert np.max(np.abs(y - want)) < 1e-3 * scale
    # the bf16 rounding is real: the f64 network differs by more than the f32 noise
    want64 = omlp.forward(p, x.astype(np.float64), n_in, hid, n_out, bf16=False)
    if hid >= 256:
        assert np.max(np.abs(want64 - want)) > 10 * np.max(np.abs(y - want))
    net.close()


@pytest.mark.parametrize("n_in,hid,n_out,rows", [(3, 256, 320, 128), (3, 48, 5, 37)])
def test_gradients_match_oracle(hip, n_in, hid, n_out, rows):
    print("mlp gradients", n_in, hid, n_out, rows, flush=True)
    net = _net(hip, n_in, hid, n_out, seed=2)
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, size=(rows, n_in)).astype(np.float32)
    t = rng.normal(size=(rows, n_out)).astype(np.float32)
    loss, g = net.gradients(x, t)
    wl, wg = omlp.loss_and_grad(p, x.astype(np.float64), t.astype(np.float64), n_in, hid, n_out, bf16=True)
    assert abs(loss - wl) < 1e-5 * max(1.0, wl)
    for name, sl in _blocks(n_in, hid, n_out):
        err = np.linalg.norm(g[sl] - wg[sl]) / np.linalg.norm(wg[sl])
        assert err < 2e-3, (name, err)
    net.close()


def test_adam_step_matches_oracle(hip):
    n_in, hid, n_out, rows = 3, 64, 16, 32
    print("mlp adam", flush=True)
    net = _net(hip, n_in, hid, n_out, seed=4)
    p0 = net.get_params().astype(np.float64)
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, size=(rows, n_in)).astype(np.float32)
    t = rng.normal(size=(rows, n_out)).astype(np.float32)
    lr = 1e-2
    _, wg = omlp.loss_and_grad(p0, x.astype(np.float64), t.astype(np.float64), n_in, hid, n_out, bf16=True)
    want, _, _ = omlp.adam_step(p0, wg, np.zeros_like(p0), np.zeros_like(p0), 1, lr)
    hist = net.train(x, t, epochs=1, batch=rows, learning_rate=lr, seed=0)
    p1 = net.get_params().astype(np.float64)
    diff = np.abs(p1 - want)
    # a first Adam step moves each weight by ~lr*sign(g); only weights whose
    # gradient is ~0 may differ in direction
    assert np.max(diff) <= 2.0 * lr + 1e-6
    assert np.mean(diff > 1e-3 * lr) < 0.01
    assert hist.shape == (1,) and np.isfinite(hist[0])
    net.close()


def test_training_reduces_loss(hip):
    print("mlp training", flush=True)
    rng = np.random.default_rng(12)
    x = rng.uniform(-1, 1, size=(1024, 3)).astype(np.float32)
    A = rng.normal(size=(3, 40))
    t = np.sin(x @ A).astype(np.float32)
    net = _net(hip, 3, 256, 40, seed=6)
    l0, _ = net.gradients(x, t)
    hist = net.train(x, t, epochs=30, batch=128, learning_rate=2e-3, seed=1)
    l1, _ = net.gradients(x, t)
    assert hist[-1] < hist[0]
    assert l1 < 0.1 * l0, (l0, l1)
    net.close()


def test_set_params_roundtrip_and_errors(hip):
    from logosq_amd import _lib
    net = _net(hip, 3, 32, 4)
    p = np.arange(net.num_params, dtype=np.float32) * 1e-4
    net.set_params(p)
    np.testing.assert_array_equal(net.get_params(), p)
    with pytest.raises(_lib.LogosQError):
        net.set_params(p[:-1])
    with pytest.raises(ValueError):
        net.forward(np.zeros((2, 5), dtype=np.float32))
    net.close()
