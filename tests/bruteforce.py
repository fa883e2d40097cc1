"""Independent numpy brute force: every gate as a full 2^n x 2^n operator
built from Kronecker products of 2x2 blocks (qubit 0 = leftmost factor,
i.e. the MSB, as src/gates/utils.rs:5-7 fixes).  It shares no index
arithmetic with the oracle's strided loops, so a wrong bit, sign or operand
order in either shows up as a mismatch."""
import math
from functools import reduce

import numpy as np

from logosq_amd import abi

I2 = np.eye(2, dtype=complex)
X2 = np.array([[0, 1], [1, 0]], dtype=complex)
Y2 = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z2 = np.array([[1, 0], [0, -1]], dtype=complex)
P0 = np.array([[1, 0], [0, 0]], dtype=complex)
P1 = np.array([[0, 0], [0, 1]], dtype=complex)
R2 = 1.0 / math.sqrt(2.0)
H2 = np.array([[R2, R2], [R2, -R2]], dtype=complex)


def op_on(n, ops):
    """Kronecker product with ops[q] on qubit q (identity elsewhere)."""
    return reduce(np.kron, [ops.get(q, I2) for q in range(n)])


def single(n, q, m):
    return op_on(n, {q: m})


def controlled(n, controls, t, m):
    """|1..1><1..1|_controls (x) m_t + (I - that projector) (x) I."""
    proj = op_on(n, {c: P1 for c in controls})
    ops = {c: P1 for c in controls}
    ops[t] = m
    full = op_on(n, ops)
    return full + (np.eye(1 << n) - proj)


def rx(th):
    c, s = math.cos(th / 2), math.sin(th / 2)
    return np.array([[c, -1j * s], [-1j * s, c]])


def ry(th):
    c, s = math.cos(th / 2), math.sin(th / 2)
    return np.array([[c, -s], [s, c]], dtype=complex)


def rz(th):
    return np.diag([np.exp(-0.5j * th), np.exp(0.5j * th)])


def u3(th, ph, la):
    c, s = math.cos(th / 2), math.sin(th / 2)
    return np.array([[c, -np.exp(1j * la) * s], [np.exp(1j * ph) * s, np.exp(1j * (ph + la)) * c]])


def two_body(n, q1, q2, th, kind):
    """exp(-i th/2 P(x)P) with the reference's RXX/RYY orientation rule: for
    q1 < q2 only the 00<->11 block rotates (parameterized.rs:563-575)."""
    P = X2 if kind == "XX" else Y2 if kind == "YY" else Z2
    PP = op_on(n, {q1: P, q2: P})
    full = math.cos(th / 2) * np.eye(1 << n) - 1j * math.sin(th / 2) * PP
    if kind == "ZZ" or q1 > q2:
        return full
    # keep the 01/10 subspace untouched
    proj_odd = op_on(n, {q1: P0, q2: P1}) + op_on(n, {q1: P1, q2: P0})
    proj_even = np.eye(1 << n) - proj_odd
    return proj_even @ full @ proj_even + proj_odd


def gate_matrix(n, g):
    k, q, th = g.kind, list(g.qubits), list(g.theta)
    if k == abi.X:
        return single(n, q[0], X2)
    if k == abi.Y:
        return single(n, q[0], -Y2)  # YGate applies -Y (single_qubit.rs:193-195)
    if k == abi.Z:
        return single(n, q[0], Z2)
    if k == abi.H:
        return single(n, q[0], H2)
    if k == abi.S:
        return single(n, q[0], np.diag([1, 1j]))
    if k == abi.T:
        return single(n, q[0], np.diag([1, np.exp(0.25j * math.pi)]))
    if k == abi.RX:
        return single(n, q[0], rx(th[0]))
    if k == abi.RY:
        return single(n, q[0], ry(th[0]))
    if k == abi.RZ:
        return single(n, q[0], rz(th[0]))
    if k == abi.PHASE:
        return single(n, q[0], np.diag([1, np.exp(1j * th[0])]))
    if k == abi.U3:
        return single(n, q[0], u3(*th))
    if k == abi.CNOT:
        return controlled(n, [q[0]], q[1], X2)
    if k == abi.CZ:
        return controlled(n, [q[0]], q[1], Z2)
    if k in (abi.CPHASE, abi.CP):
        return controlled(n, [q[0]], q[1], np.diag([1, np.exp(1j * th[0])]))
    if k == abi.CRX:
        return controlled(n, [q[0]], q[1], rx(th[0]))
    if k == abi.CRY:
        return controlled(n, [q[0]], q[1], ry(th[0]))
    if k == abi.CRZ:
        return controlled(n, [q[0]], q[1], rz(th[0]))
    if k == abi.SWAP:
        a, b = q[0], q[1]
        return sum(op_on(n, {a: pa, b: pb}) for pa, pb in
                   ((P0, P0), (P1, P1))) + op_on(n, {a: np.array([[0, 1], [0, 0]]), b: np.array([[0, 0], [1, 0]])}) \
            + op_on(n, {a: np.array([[0, 0], [1, 0]]), b: np.array([[0, 1], [0, 0]])})
    if k == abi.RXX:
        return two_body(n, q[0], q[1], th[0], "XX")
    if k == abi.RYY:
        return two_body(n, q[0], q[1], th[0], "YY")
    if k == abi.RZZ:
        return two_body(n, q[0], q[1], th[0], "ZZ")
    if k == abi.TOFFOLI:
        return controlled(n, [q[0], q[1]], q[2], X2)
    if k == abi.MATRIX1:
        m = np.array(g.matrix[:8]).view(complex).reshape(2, 2)
        return single(n, q[0], m)
    raise ValueError(k)


def pauli_string(n, codes):
    mats = {0: I2, 1: X2, 2: Y2, 3: Z2}
    return reduce(np.kron, [mats[int(c)] for c in codes])


def expectation_physical(psi, n, codes, coeffs):
    return sum(c * np.vdot(psi, pauli_string(n, row) @ psi).real for row, c in zip(codes, coeffs))


def expectation_reference(psi, n, codes, coeffs):
    """(-1)^{#Y} <P> per term: the reference's row-bit Y phase."""
    return sum(c * (-1) ** int(np.sum(np.asarray(row) == 2)) * np.vdot(psi, pauli_string(n, row) @ psi).real
               for row, c in zip(codes, coeffs))


def random_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)
