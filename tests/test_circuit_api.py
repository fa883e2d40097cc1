"""Host-side Circuit and Ansatz API against the reference's known answers
(no device): Circuit::compose / reversed (src/circuits/mod.rs:298-334,
tests/test_circuits_dense.rs:79-131), and the ansatz parameter and
operation counts (src/optimization/ansatz.rs:63-335,
tests/test_optimization_ansatz.rs:12-123)."""
import math

import pytest

import logosq_amd as lq
from logosq_amd.optimization import EntanglingGate, EntanglingPattern


def names(c):
    return [op.name() for op in c.operations()]


def test_compose_appends_in_order():
    c1 = lq.Circuit(2)
    c1.x(0)
    c2 = lq.Circuit(2)
    c2.h(1).rz(0, 0.3)
    c1.compose(c2)
    assert len(c1.operations()) == 3
    assert names(c1) == ["X", "H", "RZ"]
    assert c1.operations()[1] is c2.operations()[0]  # the same operation, shared
    assert len(c2.operations()) == 2  # the other circuit is unchanged


def test_compose_qubit_mismatch():
    c1 = lq.Circuit(2)
    with pytest.raises(lq.CircuitQubitMismatch,
                       match=r"Circuit qubit count mismatch: circuit has 2, state has 3"):
        c1.compose(lq.Circuit(3))
    assert len(c1.operations()) == 0


def test_reversed_names_and_order():
    c = lq.Circuit(1)
    c.x(0).h(0)
    r = c.reversed()
    assert len(r.operations()) == 2
    assert names(r) == ["reversed_H", "reversed_X"]
    assert [op.qubits() for op in r.operations()] == [[0], [0]]
    assert names(c) == ["X", "H"]  # the original keeps its order
    assert r.name() is None
    named = lq.Circuit(3).with_name("ladder")
    named.cnot(0, 1).ry(2, 0.5)
    rn = named.reversed()
    assert rn.name() == "ladder_reversed"
    assert [op.qubits() for op in rn.operations()] == [[2], [0, 1]]
    assert rn.operations()[0].gate is named.operations()[1].gate


def test_hardware_efficient_counts():
    a = lq.HardwareEfficientAnsatz(3, 2, EntanglingGate.CNOT, EntanglingPattern.Linear)
    assert a.num_qubits() == 3 and a.num_parameters() == 9 and a.depth == 2
    # layer 0: 3 RY + 2 CNOT; layer 1 (the last, depth > 1): 3 RY; final: 3 RY
    assert len(a.build_circuit([0.1] * 9).operations()) == 11
    lin = lq.HardwareEfficientAnsatz(4, 1, EntanglingGate.CNOT, EntanglingPattern.Linear)
    c = lin.build_circuit([0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8])
    assert c.num_qubits() == 4
    assert names(c) == ["RY"] * 4 + ["CNOT"] * 3 + ["RY"] * 4
    circ = lq.HardwareEfficientAnsatz(3, 1, EntanglingGate.CNOT, EntanglingPattern.Circular)
    assert circ.num_parameters() == 6
    assert [op.qubits() for op in circ.build_circuit([0.0] * 6).operations()[3:6]] == [[0, 1], [1, 2], [2, 0]]
    full = lq.HardwareEfficientAnsatz(3, 1, EntanglingGate.CNOT, EntanglingPattern.Full)
    c = full.build_circuit([0.1] * 6)
    assert len(c.operations()) == 9
    assert [op.qubits() for op in c.operations()[3:6]] == [[0, 1], [0, 2], [1, 2]]
    cz = lq.HardwareEfficientAnsatz(2, 1, EntanglingGate.CZ, EntanglingPattern.Circular)
    assert names(cz.build_circuit([0.0] * 4)) == ["RY", "RY", "CZ", "RY", "RY"]  # no wrap-around at n = 2


def test_hardware_efficient_wrong_parameter_count():
    a = lq.HardwareEfficientAnsatz(2, 1, EntanglingGate.CNOT, EntanglingPattern.Linear)
    with pytest.raises(AssertionError, match="Parameter count mismatch"):
        a.build_circuit([0.1, 0.2])


def test_real_amplitudes_counts():
    a = lq.RealAmplitudesAnsatz(3, 2)
    assert a.num_qubits() == 3 and a.num_parameters() == 9
    c = a.build_circuit([0.1] * 9)
    assert c.num_qubits() == 3
    assert len(c.operations()) == 2 * (3 + 2) + 3
    # depth 1 with more than one qubit adds an initial RY layer (ansatz.rs:206-212)
    one = lq.RealAmplitudesAnsatz(3, 1)
    assert one.num_parameters() == 9
    assert names(one.build_circuit([0.1] * 9)) == ["RY"] * 6 + ["CNOT"] * 2 + ["RY"] * 3
    single = lq.RealAmplitudesAnsatz(1, 1)
    assert single.num_parameters() == 2
    assert names(single.build_circuit([0.1, 0.2])) == ["RY", "RY"]


def test_efficient_su2_counts():
    a = lq.EfficientSU2Ansatz(2, 1)
    assert a.num_qubits() == 2 and a.num_parameters() == 6
    assert names(a.build_circuit([0.1] * 6)) == ["RZ", "RY", "RZ", "RZ", "RY", "RZ", "CNOT"]
    assert lq.EfficientSU2Ansatz(4, 3).num_parameters() == 36


def test_parameterized_circuit_custom():
    def builder(p):
        c = lq.Circuit(2)
        c.ry(0, p[0])
        c.ry(1, p[1])
        c.cnot(0, 1)
        c.ry(0, p[2])
        return c

    a = lq.ParameterizedCircuit(2, 3, builder)
    assert a.num_parameters() == 3 and a.num_qubits() == 2
    c = a.build_circuit([0.5, 1.0, 1.5])
    assert c.num_qubits() == 2 and names(c) == ["RY", "RY", "CNOT", "RY"]
    assert math.isclose(c.operations()[3].gate.theta[0], 1.5)
