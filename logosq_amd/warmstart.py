"""Learned warm start for VQE: a small network from Hamiltonian couplings to
ansatz parameters, trained and evaluated on the GPU.

No counterpart in the reference: its VQE::run starts from the parameters the
caller passes (src/optimization/vqe.rs:55-60) and VQE::run_random from
uniform random ones (vqe.rs:102-111).  The network here only proposes a
starting point; `choose_start` keeps it only when the simulator's f64
energy at the proposal beats the plain initialisation.
"""
import ctypes

import numpy as np

from . import _lib


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class WarmStartNet:
    """Perceptron in -> hidden -> hidden -> out, GELU (erf form), bf16
    operands with f32 accumulation on the matrix cores, f32 master weights
    and Adam on the device (lq_mlp_* in the C ABI)."""

    def __init__(self, in_dim=3, hidden=256, out_dim=320, seed=0):
        self.in_dim, self.hidden, self.out_dim = int(in_dim), int(hidden), int(out_dim)
        h = ctypes.c_void_p()
        _lib.call("lq_mlp_create", self.in_dim, self.hidden, self.out_dim, int(seed), ctypes.byref(h))
        self._h = h
        n = ctypes.c_int64()
        _lib.call("lq_mlp_num_params", self._h, ctypes.byref(n))
        self.num_params = int(n.value)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.call("lq_mlp_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        out = np.empty(self.num_params, dtype=np.float32)
        _lib.call("lq_mlp_get_params", self._h, _fptr(out), self.num_params)
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _lib.call("lq_mlp_set_params", self._h, _fptr(p), int(p.size))

    def _x(self, x):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        if x.shape[1] != self.in_dim:
            raise ValueError("inputs must have %d columns" % self.in_dim)
        return x

    def forward(self, x):
        x = self._x(x)
        y = np.empty((x.shape[0], self.out_dim), dtype=np.float32)
        _lib.call("lq_mlp_forward", self._h, _fptr(x), x.shape[0], _fptr(y))
        return y

    def gradients(self, x, target):
        x = self._x(x)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(x.shape[0], self.out_dim)
        loss = np.zeros(1, dtype=np.float32)
        g = np.empty(self.num_params, dtype=np.float32)
        _lib.call("lq_mlp_gradients", self._h, _fptr(x), _fptr(t), x.shape[0], _fptr(loss), _fptr(g))
        return float(loss[0]), g

    def train(self, x, target, epochs, batch=256, learning_rate=1e-3, seed=0):
        x = self._x(x)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(x.shape[0], self.out_dim)
        hist = np.zeros(max(int(epochs), 1), dtype=np.float32)
        _lib.call("lq_mlp_train", self._h, _fptr(x), _fptr(t), x.shape[0], int(batch), int(epochs),
                  float(learning_rate), int(seed), _fptr(hist))
        return hist[:int(epochs)]


def choose_start(energy_fn, proposal, plain):
    """Return (start, used_proposal): the proposal when its energy (f64,
    from the simulator) is below the plain initialisation's, else plain."""
    ep = energy_fn(np.asarray(proposal, dtype=np.float64))
    e0 = energy_fn(np.asarray(plain, dtype=np.float64))
    if np.isfinite(ep) and ep < e0:
        return np.asarray(proposal, dtype=np.float64).copy(), True
    return np.asarray(plain, dtype=np.float64).copy(), False
