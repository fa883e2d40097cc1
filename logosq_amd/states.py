"""Device-resident State: the mirror of `State` (src/states/mod.rs:38-392).

The amplitudes live in GPU memory owned by the C library; `vector()` copies
them to the host (the reference hands out `&Array1`, mod.rs:122-124).
"""
import ctypes
import math

import numpy as np

from . import abi
from ._lib import call, DimensionMismatch, InvalidQubitIndex, InvalidStateDimension

C128 = abi.LQ_C128
C64 = abi.LQ_C64


def _as_interleaved(vec):
    v = np.ascontiguousarray(np.asarray(vec, dtype=np.complex128))
    return v, v.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class QuantumStateBackend:
    """trait QuantumStateBackend (src/states/backend.rs:14-60): what
    Circuit.execute_on_backend drives."""

    def num_qubits(self):
        raise NotImplementedError

    def apply_single_qubit_matrix(self, qubit, matrix):
        raise NotImplementedError

    def apply_two_qubit_matrix(self, control, target, matrix):
        raise NotImplementedError

    def apply_three_qubit_matrix(self, q1, q2, q3, matrix):
        raise NotImplementedError

    def apply_full_matrix(self, matrix):
        raise NotImplementedError


def _mat_ptr(m):
    m = np.ascontiguousarray(np.asarray(m, dtype=np.complex128))
    return m, m.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class State(QuantumStateBackend):
    """A 2^n complex state vector on the GPU (complex128 or complex64)."""

    def __init__(self, num_qubits, dtype=C128):
        h = ctypes.c_void_p()
        call("lq_state_create", int(num_qubits), int(dtype), ctypes.byref(h))
        self._h = h
        self._n = int(num_qubits)
        self._dtype = int(dtype)

    # ---- constructors (mod.rs:55-170) ----
    @classmethod
    def zero_state(cls, num_qubits, dtype=C128):
        return cls(num_qubits, dtype)

    @classmethod
    def one_state(cls, num_qubits, dtype=C128):
        s = cls(num_qubits, dtype)
        call("lq_state_set_basis", s._h, (1 << num_qubits) - 1)
        return s

    @classmethod
    def plus_state(cls, num_qubits, dtype=C128):
        s = cls(num_qubits, dtype)
        call("lq_state_set_plus", s._h)
        return s

    @classmethod
    def bell_state(cls, dtype=C128):
        r = 1.0 / math.sqrt(2.0)
        s = cls(2, dtype)
        s._upload(np.array([r, 0, 0, r], dtype=np.complex128), normalize=False)
        return s

    @classmethod
    def new(cls, vector, num_qubits=None, dtype=C128):
        """State::new (mod.rs:55-75): validate the power-of-two length, normalise."""
        vec = np.asarray(vector, dtype=np.complex128)
        n = num_qubits if num_qubits is not None else int(math.log2(len(vec))) if len(vec) > 0 else 0
        if len(vec) != (1 << n) or n < 1:
            raise InvalidStateDimension(abi.LQ_ERR_INVALID_STATE_DIMENSION,
                                        "Invalid state dimension: %d (must be a power of 2)" % len(vec))
        s = cls(n, dtype)
        s._upload(vec, normalize=True)
        return s

    @classmethod
    def new_unchecked(cls, vector, num_qubits, dtype=C128):
        s = cls(num_qubits, dtype)
        s._upload(np.asarray(vector, dtype=np.complex128), normalize=True)
        return s

    def _upload(self, vec, normalize):
        v, p = _as_interleaved(vec)
        call("lq_state_upload", self._h, p, len(v), 1 if normalize else 0)

    # ---- accessors ----
    @property
    def handle(self):
        return self._h

    @property
    def dtype(self):
        return self._dtype

    def num_qubits(self):
        return self._n

    def vector(self):
        out = np.empty(1 << self._n, dtype=np.complex128)
        call("lq_state_download", self._h, out.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
             len(out))
        return out

    def set_vector(self, vec):
        """Raw write through vector_mut (mod.rs:131-133): no normalisation."""
        self._upload(vec, normalize=False)

    def normalize(self):
        norm = ctypes.c_double()
        call("lq_state_normalize", self._h, ctypes.byref(norm))
        return norm.value

    def norm_sqr(self):
        out = ctypes.c_double()
        call("lq_state_norm_sqr", self._h, ctypes.byref(out))
        return out.value

    def probabilities(self):
        out = np.empty(1 << self._n, dtype=np.float64)
        call("lq_state_probabilities", self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(out))
        return out

    def probability(self, basis_state):
        """probability (mod.rs:288-293): 0.0 when out of range."""
        if basis_state >= (1 << self._n):
            return 0.0
        return float(abs(self.vector()[basis_state]) ** 2)

    # ---- measurement (mod.rs:180-262); the reference's RNG is unseeded, so
    # these take uniforms or a seed ----
    def bit_probabilities(self):
        """p1[b] = P(bit b of the basis index is 1)."""
        out = np.empty(self._n, dtype=np.float64)
        call("lq_bit_probabilities", self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def measure_qubit(self, qubit_idx, u=None):
        """State::measure_qubit (mod.rs:196-231): mask 1 << qubit_idx, outcome
        1 when u < P(1), then project and normalise."""
        if u is None:
            u = float(np.random.random())
        r = ctypes.c_int()
        call("lq_measure_qubit", self._h, int(qubit_idx), float(u), ctypes.byref(r))
        return r.value

    def sample(self, shots, seed=0):
        """`shots` draws of State::measure (no collapse) as basis indices."""
        out = np.empty(int(shots), dtype=np.uint64)
        call("lq_sample_shots", self._h, int(seed) & ((1 << 64) - 1), int(shots),
             out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        return out

    def measure(self, seed=None):
        """State::measure (mod.rs:180-185): one basis index drawn from |a|^2."""
        if seed is None:
            seed = int(np.random.randint(0, 2 ** 63))
        return int(self.sample(1, seed)[0])

    def measure_shots(self, n_shots, seed=None):
        """measure_shots_parallel / measure_shots (mod.rs:241-262): counts."""
        if seed is None:
            seed = int(np.random.randint(0, 2 ** 63))
        idx, cnt = np.unique(self.sample(n_shots, seed), return_counts=True)
        return {int(i): int(c) for i, c in zip(idx, cnt)}

    measure_shots_parallel = measure_shots

    def inner_product(self, other):
        out = (ctypes.c_double * 2)()
        call("lq_state_inner_product", self._h, other._h, out)
        return complex(out[0], out[1])

    inner_product_parallel = inner_product

    def clone(self):
        s = State(self._n, self._dtype)
        call("lq_state_copy", s._h, self._h)
        return s

    def apply_gate(self, gate):
        gate.apply(self)

    # ---- QuantumStateBackend for the dense state (src/states/mod.rs:394-487):
    # each method is the reference's expand-then-apply_full_matrix, i.e. the
    # block on its qubits followed by State::normalize, without building the
    # 2^n x 2^n matrix ----
    def apply_single_qubit_matrix(self, qubit, matrix):
        if qubit < 0 or qubit >= self._n:
            raise InvalidQubitIndex(abi.LQ_ERR_INVALID_QUBIT_INDEX,
                                    "Invalid qubit index: %d (must be < %d)" % (qubit, self._n))
        m, p = _mat_ptr(np.asarray(matrix).reshape(2, 2))
        call("lq_apply_1q", self._h, int(qubit), p)
        self.normalize()

    def apply_two_qubit_matrix(self, control, target, matrix):
        if max(control, target) >= self._n or min(control, target) < 0:
            raise InvalidQubitIndex(abi.LQ_ERR_INVALID_QUBIT_INDEX,
                                    "Invalid qubit index: %d (must be < %d)" % (max(control, target), self._n))
        m, p = _mat_ptr(np.asarray(matrix).reshape(4, 4))
        call("lq_apply_2q", self._h, int(control), int(target), p)
        self.normalize()

    def apply_three_qubit_matrix(self, q1, q2, q3, matrix):
        # the reference applies `matrix` as the full-state operator (mod.rs:459-469)
        self.apply_full_matrix(matrix)

    def apply_full_matrix(self, matrix):
        m = np.asarray(matrix)
        dim = 1 << self._n
        if m.ndim != 2 or m.shape[1] != dim or m.shape[0] != dim:
            raise DimensionMismatch(abi.LQ_ERR_DIMENSION_MISMATCH, "Dimension mismatch: expected %d, got %d"
                                    % (dim, m.shape[1] if m.ndim == 2 else -1))
        m, p = _mat_ptr(m)
        call("lq_apply_dense", self._h, p, dim)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                from ._lib import load
                load().lq_state_destroy(h)
            except Exception:
                pass
            self._h = ctypes.c_void_p()
