"""Circuit::expectation(&dyn Gate) (src/circuits/mod.rs:256-287) and the
executed forms of compose / reversed and the ansatzes on the device.

Circuit::expectation re-executes the circuit from |0...0>, applies the
observable gate to a copy and returns Re <psi|O psi>.  Known answers from
tests/test_circuits_dense.rs:778-797 (<+|X|+> = 1, <+|Z|+> = 0, with the
dense X / Z matrix gates), then random circuits with gates of every family as
the observable against the oracle: execute, apply_gate (or the MatrixGate
loop with its renormalisation) and inner_product, <= 1e-12."""
import math

import numpy as np
import pytest

import oracle
import logosq_amd as lq
from logosq_amd import abi
from tests import bruteforce as bf
from tests.test_planner import random_circuit

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_expectation_known_answers(hip):
    say("Circuit::expectation <+|X|+>, <+|Z|+>")
    c = lq.Circuit(1)
    c.h(0)
    assert abs(c.expectation(lq.x_gate()) - 1.0) < 1e-10
    assert abs(c.expectation(lq.z_gate())) < 1e-10
    # the struct gates give the same on |+>; the reference's YGate applies -Y
    assert abs(c.expectation(lq.gates.x(0, 1)) - 1.0) < 1e-12
    assert abs(c.expectation(lq.gates.y(0, 1))) < 1e-12
    # |0>: <Z> = 1, <X> = 0
    assert abs(lq.Circuit(1).expectation(lq.z_gate()) - 1.0) < 1e-12
    assert abs(lq.Circuit(1).expectation(lq.x_gate())) < 1e-12


def _observables(n, rng):
    th = rng.uniform(0, 2 * math.pi, 4)
    return [
        lq.gates.z(n - 1, n), lq.gates.x(0, n), lq.gates.y(n // 2, n), lq.gates.h(1, n),
        lq.gates.rx(2, th[0], n), lq.gates.rz(0, th[1], n), lq.gates.cnot(0, n - 1, n),
        lq.gates.cz(1, 3, n), lq.gates.swap(0, 2, n), lq.gates.crx(n - 1, 1, th[2], n),
        lq.gates.rxx(2, 1, th[3], n), lq.gates.toffoli(0, 1, n - 1, n),
    ]


@pytest.mark.parametrize("n", [4, 7, 11])
def test_expectation_random_circuits(hip, n):
    say("Circuit::expectation random n=%d" % n)
    rng = np.random.default_rng(8100 + n)
    gates = random_circuit(n, 60, rng, with_dense=False)
    c = lq.Circuit(n)
    for g in gates:
        c.add_operation(g, [int(q) for q in g.qubits[: abi.ARITY[g.kind]]], "op")
    psi = oracle.execute(oracle.zero_state(n), n, gates)
    for obs in _observables(n, rng):
        want = oracle.inner_product(psi, oracle.apply_gate(psi, n, obs.desc())).real
        got = c.expectation(obs)
        assert abs(got - want) <= 1e-12, (type(obs).__name__, got, want)
    # a MatrixGate observable: dense Hermitian matrix; the gate renormalises O psi
    m = bf.random_state(2 * n, rng).reshape(1 << n, 1 << n)
    m = m + m.conj().T
    want = oracle.inner_product(psi, oracle.matrix_gate(psi, n, m)).real
    assert abs(c.expectation(lq.MatrixGate(m)) - want) <= 1e-12


def test_composed_and_reversed_execution(hip):
    """tests/test_circuits_dense.rs:79-131: X(0) composed with H(1) gives
    (|10> + |11>)/sqrt(2); X -> H reversed is H -> X, both with
    probabilities 1/2, 1/2."""
    say("compose / reversed execution")
    c1 = lq.Circuit(2)
    c1.x(0)
    c2 = lq.Circuit(2)
    c2.h(1)
    c1.compose(c2)
    st = lq.State.zero_state(2)
    c1.execute(st)
    p = st.probabilities()
    assert p[0] < 1e-10 and p[1] < 1e-10 and abs(p[2] - 0.5) < 1e-10 and abs(p[3] - 0.5) < 1e-10
    c = lq.Circuit(1)
    c.x(0).h(0)
    s1, s2 = lq.State.zero_state(1), lq.State.zero_state(1)
    c.execute(s1)
    c.reversed().execute(s2)
    r = 1.0 / math.sqrt(2.0)
    assert np.max(np.abs(s1.vector() - np.array([r, -r]))) < 1e-15
    assert np.max(np.abs(s2.vector() - np.array([r, r]))) < 1e-15


def test_ansatz_apply_normalised(hip):
    """tests/test_optimization_ansatz.rs:54-96: HEA(2, 1) at pi/2 and
    EfficientSU2(2, 1) at 0.1 leave a normalised state, equal to the oracle's."""
    say("ansatz apply")
    for ans, x in ((lq.HardwareEfficientAnsatz(2, 1), [math.pi / 2] * 4), (lq.EfficientSU2Ansatz(2, 1), [0.1] * 6),
                   (lq.RealAmplitudesAnsatz(3, 1), list(np.linspace(0.1, 0.9, 9)))):
        st = lq.State.zero_state(ans.num_qubits())
        ans.apply(st, x)
        assert abs(float(np.sum(st.probabilities())) - 1.0) < 1e-10
        n = ans.num_qubits()
        want = oracle.execute(oracle.zero_state(n), n, ans.build_circuit(x).descriptors())
        assert np.max(np.abs(st.vector() - want)) <= 1e-12
