"""Circuit: the mirror of `Circuit` (src/circuits/mod.rs:81-375 and the
builder impls in src/circuits/{single_qubit,two_qubits,three_qubits,
parameterized}.rs).

Operations are recorded as lq_gate descriptors; `execute` hands the whole
list to the library (lq_execute), which runs it either gate by gate or in
fused LDS tile passes.  Builder panics of the reference become exceptions
with the same message.
"""
import ctypes
import math

import numpy as np

from . import abi
from ._lib import call, CircuitQubitMismatch, InvalidQubitIndex
from .gates import Gate, MatrixGate


class Param:
    """A parameter slot for ansatz templates: pass Param(i, value) where a
    builder expects an angle to record `param = i` in the descriptor."""

    __slots__ = ("index", "value")

    def __init__(self, index, value=0.0):
        self.index = int(index)
        self.value = float(value)

    def __float__(self):
        return self.value


def _angle(a):
    if isinstance(a, Param):
        return a.value, a.index
    return float(a), -1


class Operation:
    """Operation (src/circuits/mod.rs:29-59)."""

    def __init__(self, gate, qubits, name):
        self.gate = gate          # lq_gate descriptor or a Gate object
        self._qubits = list(qubits)
        self._name = name

    def qubits(self):
        return self._qubits

    def name(self):
        return self._name

    def __repr__(self):
        return "%s(qubits: %s)" % (self._name, self._qubits)


def _local_matrix(gate, qubits):
    """The gate's matrix on its operation's k qubits (sub-index: bit of
    qubits[0] first), from its action on a k-qubit device state whose qubits
    keep the order of the originals (the reference's quirks depend on it)."""
    from .states import State
    k = len(qubits)
    local = {q: r for r, q in enumerate(sorted(qubits))}
    g = abi.LqGate()
    ctypes.memmove(ctypes.addressof(g), ctypes.addressof(gate), ctypes.sizeof(gate))
    for i in range(3):
        if g.qubits[i] >= 0:
            g.qubits[i] = local[g.qubits[i]]

    def basis(sub):
        return sum(((sub >> (k - 1 - i)) & 1) << (k - 1 - local[qubits[i]]) for i in range(k))

    st = State(k)
    m = np.zeros((1 << k, 1 << k), dtype=np.complex128)
    for col in range(1 << k):
        call("lq_state_set_basis", st.handle, basis(col))
        call("lq_apply_gate", st.handle, ctypes.byref(g))
        v = st.vector()
        for row in range(1 << k):
            m[row, col] = v[basis(row)]
    return m


def _full_matrix(op, n):
    """An operation's full 2^n matrix probed on basis states (mod.rs:456-467)."""
    from .states import State
    dim = 1 << n
    m = np.zeros((dim, dim), dtype=np.complex128)
    st = State(n)
    for col in range(dim):
        call("lq_state_set_basis", st.handle, col)
        if isinstance(op.gate, abi.LqGate):
            call("lq_apply_gate", st.handle, ctypes.byref(op.gate))
        else:
            op.gate.apply(st)
        m[:, col] = st.vector()
    return m


class Circuit:
    def __init__(self, num_qubits):
        self._n = int(num_qubits)
        self._ops = []
        self._name = None
        self._noise = []

    def add_noise(self, noise_model):
        """Circuit::add_noise (mod.rs:339-342): applied after every
        operation by execute."""
        self._noise.append(noise_model)
        return self

    def with_name(self, name):
        self._name = name
        return self

    def num_qubits(self):
        return self._n

    def num_operations(self):
        return len(self._ops)

    def operations(self):
        return self._ops

    def name(self):
        return self._name

    # ---- generic adders (mod.rs:137-187) ----
    def add_operation(self, gate, qubits, name):
        for q in qubits:
            if q >= self._n:
                raise InvalidQubitIndex(abi.LQ_ERR_INVALID_QUBIT_INDEX,
                                        "Invalid qubit index: %d (must be < %d)" % (q, self._n))
        self._ops.append(Operation(gate, qubits, name))

    def add_matrix_gate(self, matrix, qubits, name):
        self.add_operation(MatrixGate(matrix), qubits, name)

    def _desc(self, kind, qubits, angles=(), name="", flags=0, matrix=None, op_qubits=None):
        th, pr = [], []
        for a in angles:
            v, i = _angle(a)
            th.append(v)
            pr.append(i)
        g = abi.make_gate(kind, qubits, th, pr, flags, matrix)
        self.add_operation(g, op_qubits if op_qubits is not None else list(qubits), name)
        return self

    # ---- dense builders (single_qubit.rs:12-126) ----
    def add_single_qubit_gate(self, gate_matrix, target_qubit, name):
        m = np.asarray(gate_matrix, dtype=np.complex128)
        if m.shape != (2, 2):
            raise AssertionError("Gate must be a 2x2 matrix for single-qubit operation")
        if not target_qubit < self._n:
            raise AssertionError("Target qubit index out of range")
        return self._desc(abi.MATRIX1, [target_qubit], (), name, abi.GATE_FLAG_RENORMALIZE, m.reshape(-1),
                          op_qubits=list(range(self._n)))

    def x(self, q):
        return self.add_single_qubit_gate([[0, 1], [1, 0]], q, "X")

    def y(self, q):
        return self.add_single_qubit_gate([[0, -1j], [1j, 0]], q, "Y")

    def z(self, q):
        return self.add_single_qubit_gate([[1, 0], [0, -1]], q, "Z")

    def h(self, q):
        r = 1.0 / math.sqrt(2.0)
        return self.add_single_qubit_gate([[r, r], [r, -r]], q, "H")

    def s(self, q):
        return self.add_single_qubit_gate([[1, 0], [0, 1j]], q, "S")

    def t(self, q):
        r = 1.0 / math.sqrt(2.0)
        return self.add_single_qubit_gate([[1, 0], [0, complex(r, r)]], q, "T")

    def rx_matrix(self, q, theta):
        c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
        return self.add_single_qubit_gate([[c, -1j * s], [-1j * s, c]], q, "RX(%s)" % theta)

    def ry_matrix(self, q, theta):
        c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
        return self.add_single_qubit_gate([[c, -s], [s, c]], q, "RY(%s)" % theta)

    def rz_matrix(self, q, theta):
        m = [[complex(math.cos(theta / 2.0), -math.sin(theta / 2.0)), 0],
             [0, complex(math.cos(theta / 2.0), math.sin(theta / 2.0))]]
        return self.add_single_qubit_gate(m, q, "RZ(%s)" % theta)

    # ---- two- and three-qubit builders (two_qubits.rs, three_qubits.rs) ----
    def _check2(self, a, b, same_msg):
        if not (a < self._n and b < self._n):
            raise AssertionError("Qubit indices out of range")
        if a == b:
            raise AssertionError(same_msg)

    def add_two_qubit_gate(self, gate_matrix, qubit1, qubit2, name):
        m = np.asarray(gate_matrix, dtype=np.complex128)
        if m.shape != (4, 4):
            raise AssertionError("Gate must be a 4x4 matrix for two-qubit operation")
        self._check2(qubit1, qubit2, "Qubits must be different")
        q1, q2 = min(qubit1, qubit2), max(qubit1, qubit2)
        oq = [q1, q2] if self._n == 2 else list(range(self._n))
        return self._desc(abi.MATRIX2, [q1, q2], (), name, abi.GATE_FLAG_RENORMALIZE, m.reshape(-1), op_qubits=oq)

    def cnot(self, c, t):
        self._check2(c, t, "Control and target must be different")
        return self._desc(abi.CNOT, [c, t], (), "CNOT")

    def swap(self, a, b):
        self._check2(a, b, "Qubits must be different")
        return self._desc(abi.SWAP, [a, b], (), "SWAP")

    def cz(self, c, t):
        self._check2(c, t, "Control and target must be different")
        return self._desc(abi.CZ, [c, t], (), "CZ")

    def toffoli(self, c1, c2, t):
        if not (c1 < self._n and c2 < self._n and t < self._n):
            raise AssertionError("Qubit indices out of range")
        if c1 == c2 or c1 == t or c2 == t:
            raise AssertionError("Control and target qubits must be different")
        return self._desc(abi.TOFFOLI, [c1, c2, t], (), "Toffoli")

    # ---- parameterised builders (parameterized.rs) ----
    def _check1(self, q):
        if not q < self._n:
            raise AssertionError("Qubit index out of range")

    def rx(self, q, theta):
        self._check1(q)
        return self._desc(abi.RX, [q], (theta,), "RX")

    def ry(self, q, theta):
        self._check1(q)
        return self._desc(abi.RY, [q], (theta,), "RY")

    def rz(self, q, theta):
        self._check1(q)
        return self._desc(abi.RZ, [q], (theta,), "RZ")

    def phase(self, q, theta):
        self._check1(q)
        return self._desc(abi.PHASE, [q], (theta,), "P")

    def u3(self, q, theta, phi, lam):
        self._check1(q)
        return self._desc(abi.U3, [q], (theta, phi, lam), "U3")

    def crx(self, c, t, theta):
        self._check2(c, t, "Control and target must be different")
        return self._desc(abi.CRX, [c, t], (theta,), "CRX")

    def cry(self, c, t, theta):
        self._check2(c, t, "Control and target must be different")
        return self._desc(abi.CRY, [c, t], (theta,), "CRY")

    def crz(self, c, t, theta):
        self._check2(c, t, "Control and target must be different")
        return self._desc(abi.CRZ, [c, t], (theta,), "CRZ")

    def cphase_param(self, c, t, theta):
        self._check2(c, t, "Control and target must be different")
        return self._desc(abi.CP, [c, t], (theta,), "CP")

    def rxx(self, a, b, theta):
        self._check2(a, b, "Qubits must be different")
        return self._desc(abi.RXX, [a, b], (theta,), "RXX")

    def ryy(self, a, b, theta):
        self._check2(a, b, "Qubits must be different")
        return self._desc(abi.RYY, [a, b], (theta,), "RYY")

    def rzz(self, a, b, theta):
        self._check2(a, b, "Qubits must be different")
        return self._desc(abi.RZZ, [a, b], (theta,), "RZZ")

    def controlled_phase(self, c, t, angle):
        """qft::controlled_phase (src/algorithms/qft.rs:96-105): a CPhaseGate."""
        return self._desc(abi.CPHASE, [c, t], (angle,), "CP(%.4f)" % angle)

    # ---- composition (mod.rs:298-334) ----
    def compose(self, other):
        if self._n != other._n:
            raise CircuitQubitMismatch(abi.LQ_ERR_CIRCUIT_QUBIT_MISMATCH,
                                       "Circuit qubit count mismatch: circuit has %d, state has %d"
                                       % (self._n, other._n))
        self._ops.extend(other._ops)

    def reversed(self):
        r = Circuit(self._n)
        for op in reversed(self._ops):
            r._ops.append(Operation(op.gate, op.qubits(), "reversed_" + op.name()))
        if self._name is not None:
            r._name = self._name + "_reversed"
        return r

    # ---- execution (mod.rs:196-209, 352-375) ----
    def descriptors(self):
        """The recorded lq_gate descriptors (None for dense MatrixGate ops)."""
        return [op.gate if isinstance(op.gate, abi.LqGate) else None for op in self._ops]

    def execute(self, state, mode=None):
        """Circuit::execute (mod.rs:352-375): every operation in order, then
        every noise model after each operation when there are any."""
        if not self._noise:
            return self.execute_without_noise(state, mode)
        self._check_state(state)
        for op in self._ops:
            if isinstance(op.gate, abi.LqGate):
                call("lq_apply_gate", state.handle, ctypes.byref(op.gate))
            else:
                op.gate.apply(state)
            for model in self._noise:
                model.apply(state)

    def _check_state(self, state):
        if state.num_qubits() != self._n:
            raise CircuitQubitMismatch(abi.LQ_ERR_CIRCUIT_QUBIT_MISMATCH,
                                       "Circuit qubit count mismatch: circuit has %d, state has %d"
                                       % (self._n, state.num_qubits()))

    def execute_without_noise(self, state, mode=None):
        """Circuit::execute_without_noise (mod.rs:196-209); also execute
        when no noise model is attached."""
        self._check_state(state)
        if mode is None:
            mode = default_exec_mode()
        run = []
        for op in self._ops:
            if isinstance(op.gate, abi.LqGate):
                run.append(op.gate)
                continue
            self._flush(state, run, mode)
            run = []
            op.gate.apply(state)
        self._flush(state, run, mode)

    def execute_and_measure(self, uniforms=None):
        """Circuit::execute_and_measure (mod.rs:215-233, the default parallel
        build): run from |0...0>, then measure_qubit(i) for i = 0..n-1, each on
        its own copy of the final state, i.e. bit i against P(bit i = 1).
        `uniforms` (n values in [0, 1)) replace the reference's thread_rng."""
        from .states import State
        st = State.zero_state(self._n)
        self.execute(st)
        p1 = st.bit_probabilities()
        u = np.random.random(self._n) if uniforms is None else np.asarray(uniforms, dtype=np.float64)
        return [1 if u[i] < p1[i] else 0 for i in range(self._n)]

    def execute_on_backend(self, backend):
        """Circuit::execute_on_backend (mod.rs:444-504) with descriptor
        dispatch.  The reference probes every operation's full 2^n matrix on
        basis states (mod.rs:456-467) and reads from it the block on the
        operation's qubits with every other qubit 0 (mod.rs:508-590); here
        that block is read from the gate itself, applied to a state of as many
        qubits as the operation has.  Three-or-more-qubit operations go to
        apply_three_qubit_matrix / apply_full_matrix with the full matrix, as
        in the reference; on the device State that full-state operator is
        applied as the gate followed by State::normalize."""
        from .states import State
        if backend.num_qubits() != self._n:
            raise CircuitQubitMismatch(abi.LQ_ERR_CIRCUIT_QUBIT_MISMATCH,
                                       "Circuit qubit count mismatch: circuit has %d, state has %d"
                                       % (self._n, backend.num_qubits()))
        for op in self._ops:
            qs = op.qubits()
            if len(qs) in (1, 2) and isinstance(op.gate, abi.LqGate):
                m = _local_matrix(op.gate, qs)
                if len(qs) == 1:
                    backend.apply_single_qubit_matrix(qs[0], m)
                else:
                    backend.apply_two_qubit_matrix(qs[0], qs[1], m)
                continue
            if isinstance(backend, State):
                if isinstance(op.gate, abi.LqGate):
                    call("lq_apply_gate", backend.handle, ctypes.byref(op.gate))
                else:
                    op.gate.apply(backend)
                backend.normalize()
                continue
            full = _full_matrix(op, self._n)
            if len(qs) == 3:
                backend.apply_three_qubit_matrix(qs[0], qs[1], qs[2], full)
            else:
                backend.apply_full_matrix(full)

    @staticmethod
    def _flush(state, run, mode):
        if not run:
            return
        arr = abi.gate_array(run)
        call("lq_execute", state.handle, arr, len(run), mode)

    def expectation(self, observable):
        """Circuit::expectation (mod.rs:256-287): Re <psi|O|psi> with O a Gate."""
        from .states import State
        st = State.zero_state(self._n)
        self.execute(st)
        obs = st.clone()
        observable.apply(obs)
        return st.inner_product(obs).real


_EXEC_MODE = [abi.EXEC_PER_GATE]


def default_exec_mode():
    return _EXEC_MODE[0]


def set_default_exec_mode(mode):
    _EXEC_MODE[0] = mode
