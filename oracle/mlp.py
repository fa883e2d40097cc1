"""CPU oracle of the warm-start network (float64 numpy).

TEST INFRASTRUCTURE ONLY: used by tests/ and bench.py as the checker of the
GPU training and inference kernels; the product path never imports it.

The network has no counterpart in the reference (the reference's VQE starts
from caller-given parameters, src/optimization/vqe.rs:55-60).  It is
in -> H -> H -> out with the erf GELU after both hidden layers; parameters
are one vector [W1 (H x in), b1, W2 (H x H), b2, W3 (out x H), b3], row-major.
With ``bf16=True`` every matrix-product operand is first rounded to f32 and
then to bf16 (round to nearest even), as the GPU does; sums stay float64.
The loss is the mean squared error over rows x out.
"""
import numpy as np
from scipy.special import erf


def bf16_round(a):
    """float -> nearest bf16 value (ties to even), via f32, as float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, u | 0x400000, r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def num_params(n_in, hid, n_out):
    return hid * n_in + hid + hid * hid + hid + n_out * hid + n_out


def unpack(p, n_in, hid, n_out):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    out = []
    for shape in ((hid, n_in), (hid,), (hid, hid), (hid,), (n_out, hid), (n_out,)):
        size = int(np.prod(shape))
        out.append(p[o:o + size].reshape(shape))
        o += size
    assert o == p.size, "parameter vector has the wrong length"
    return out


def forward(p, x, n_in, hid, n_out, bf16=True, keep=False):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, hid, n_out)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    x = np.asarray(x, dtype=np.float64)
    z1 = r(x) @ r(W1).T + b1
    a1 = gelu(z1)
    z2 = r(a1) @ r(W2).T + b2
    a2 = gelu(z2)
    y = r(a2) @ r(W3).T + b3
    if keep:
        return y, (x, z1, a1, z2, a2)
    return y


def loss_and_grad(p, x, t, n_in, hid, n_out, bf16=True):
    """Mean squared error and its gradient with respect to p (same layout)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, hid, n_out)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, (x, z1, a1, z2, a2) = forward(p, x, n_in, hid, n_out, bf16, keep=True)
    t = np.asarray(t, dtype=np.float64)
    e = y - t
    loss = float(np.mean(e * e))
    d3 = 2.0 * e / e.size
    gW3 = r(d3).T @ r(a2)
    gb3 = d3.sum(axis=0)
    d2 = (r(d3) @ r(W3)) * gelu_grad(z2)
    gW2 = r(d2).T @ r(a1)
    gb2 = d2.sum(axis=0)
    d1 = (r(d2) @ r(W2)) * gelu_grad(z1)
    gW1 = r(d1).T @ r(x)
    gb1 = d1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns new (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
