"""Circuit.execute_on_backend on the device State (QuantumStateBackend,
src/states/backend.rs) against the oracle's restatement of the reference's
backend path (probe, extract, expand, dense apply + normalize): amplitudes
<= 1e-12.  A host-side backend driven by the same dispatch receives the
same matrices."""
import math

import numpy as np
import pytest

import logosq_amd as lq
from oracle import backend as ob

pytestmark = pytest.mark.gpu


def _circuit(n):
    c = lq.Circuit(n)
    c.h(0)  # dense builder: operation on every qubit
    c.rx(1, 0.3)
    c.cnot(2, 0)
    c.rxx(3, 1, 0.9)
    c.toffoli(0, 1, 3)
    c.u3(2, 0.2, 0.4, 0.6)
    c.cry(1, 4, 1.1)
    c.swap(0, 4)
    c.y(1)
    c.rz(4, 2.2)
    c.cz(3, 2)
    return c


class _HostBackend(lq.states.QuantumStateBackend):
    """A numpy dense backend with the reference State's semantics."""

    def __init__(self, v, n):
        self.v, self.n = np.array(v, dtype=np.complex128), n

    def num_qubits(self):
        return self.n

    def _full(self, m):
        self.v = m @ self.v
        self.v /= np.linalg.norm(self.v)

    def apply_single_qubit_matrix(self, qubit, matrix):
        import oracle
        self._full(oracle.kron_single(np.asarray(matrix), self.n, qubit))

    def apply_two_qubit_matrix(self, control, target, matrix):
        self._full(ob.expand_two_backend(np.asarray(matrix), self.n, control, target))

    def apply_three_qubit_matrix(self, q1, q2, q3, matrix):
        self._full(np.asarray(matrix))

    def apply_full_matrix(self, matrix):
        self._full(np.asarray(matrix))


def test_execute_on_backend_matches_oracle(hip):
    n = 5
    print("backend path n=%d" % n, flush=True)
    rng = np.random.default_rng(17)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    c = _circuit(n)
    st = lq.State.new(v)
    c.execute_on_backend(st)
    want = ob.execute_on_backend(v / np.linalg.norm(v), n, c.descriptors(), [op.qubits() for op in c.operations()])
    assert np.max(np.abs(st.vector() - want)) <= 1e-12
    hb = _HostBackend(v / np.linalg.norm(v), n)
    c.execute_on_backend(hb)
    assert np.max(np.abs(hb.v - want)) <= 1e-12


def test_backend_errors(hip):
    st = lq.State(3)
    with pytest.raises(lq.InvalidQubitIndex):
        st.apply_single_qubit_matrix(3, np.eye(2))
    with pytest.raises(lq.DimensionMismatch):
        st.apply_full_matrix(np.eye(4))
    with pytest.raises(lq.CircuitQubitMismatch):
        lq.Circuit(4).execute_on_backend(st)
    # the single-qubit backend call renormalises, as the reference's dense apply does
    st.apply_single_qubit_matrix(0, 2.0 * np.eye(2))
    assert abs(st.norm_sqr() - 1.0) < 1e-15
    assert math.isclose(abs(st.vector()[0]), 1.0)
