"""Pins the CPU restatement of the reference's QuantumStateBackend path
(oracle/backend.py) without a GPU: the two-qubit expansion against the
circuit builder's sorted-pair expansion, block extraction against closed
forms, and the whole backend execution against direct gate application (the
same operator for unitary gates, up to rounding)."""
import math

import numpy as np

import oracle
from oracle import backend as ob
from logosq_amd import abi


def test_expand_two_matches_sorted_pair_builder():
    rng = np.random.default_rng(3)
    m = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    n = 4
    for c, t in ((0, 2), (1, 3), (0, 1)):
        np.testing.assert_array_equal(ob.expand_two_backend(m, n, c, t), oracle.expand_two(m, n, c, t))
    # reversed order: the sub-index bits swap
    perm = [0, 2, 1, 3]
    mp = m[np.ix_(perm, perm)]
    np.testing.assert_array_equal(ob.expand_two_backend(m, n, 3, 1), oracle.expand_two(mp, n, 1, 3))


def test_extract_single_closed_form():
    n, q, th = 3, 1, 0.7
    full = ob.full_matrix(n, abi.make_gate(abi.RX, [q], [th]))
    c, s = math.cos(th / 2), math.sin(th / 2)
    np.testing.assert_allclose(ob.extract_single(full, q, n), [[c, -1j * s], [-1j * s, c]], atol=1e-15)
    # the reference's YGate applies -Y: its block is [[0, i], [-i, 0]]
    full = ob.full_matrix(n, abi.make_gate(abi.Y, [2]))
    np.testing.assert_allclose(ob.extract_single(full, 2, n), [[0, 1j], [-1j, 0]], atol=0)


def test_backend_path_equals_direct_application():
    n = 4
    rng = np.random.default_rng(5)
    gates = [abi.make_gate(abi.H, [0]), abi.make_gate(abi.RY, [1], [0.3]), abi.make_gate(abi.CNOT, [2, 0]),
             abi.make_gate(abi.RXX, [3, 1], [0.9]), abi.make_gate(abi.TOFFOLI, [0, 1, 3]),
             abi.make_gate(abi.U3, [2], [0.2, 0.4, 0.6]), abi.make_gate(abi.CRY, [1, 3], [1.1]),
             abi.make_gate(abi.SWAP, [0, 3]), abi.make_gate(abi.Y, [1])]
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    v /= np.linalg.norm(v)
    a = ob.execute_on_backend(v, n, gates)
    b = oracle.execute(v, n, gates)
    assert np.max(np.abs(a - b)) < 1e-13
