"""Writes tests/golden/fixtures.npz: seeded inputs and the oracle's outputs,
each cross-checked here against an independent computation before it is
written (numpy Kronecker brute force, the QFT closed form sqrt(N) * ifft).

    python tests/golden/make_golden.py

Contents (all complex arrays complex128):
  qft12_in, qft12_out        QFT on 12 qubits of a seeded N(0,1) state
                             (normalised), the gate list of qft::create_circuit
  gates_n, gates_in          a seeded 5-qubit state
  gates_desc                 int/float description of every (gate, qubits,
                             angle) case: kind, q0, q1, q2, theta0..2
  gates_out                  the state after each single gate
  pauli_codes, pauli_coeffs  20 random 6-qubit Pauli sums (with Y terms)
  pauli_state                a seeded 6-qubit state
  pauli_ref, pauli_phys      their expectations, reference and physical Y
  ps_params, ps_energy, ps_grad, ps_codes, ps_coeffs
                             parameter shift of the 8-qubit, L=2 RY/RZ+CZ ring
                             ansatz on an XYZ ring
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
import bruteforce as bf  # noqa: E402
from logosq_amd import abi, algorithms  # noqa: E402


def gate_cases(n):
    one = [abi.X, abi.Y, abi.Z, abi.H, abi.S, abi.T]
    rot = [abi.RX, abi.RY, abi.RZ, abi.PHASE]
    two = [abi.CNOT, abi.SWAP, abi.CZ]
    two_rot = [abi.CPHASE, abi.CRX, abi.CRY, abi.CRZ, abi.CP, abi.RXX, abi.RYY, abi.RZZ]
    cases = []
    for q in range(n):
        for k in one:
            cases.append((k, [q], []))
        for k in rot:
            cases.append((k, [q], [0.37 + 0.5 * q]))
        cases.append((abi.U3, [q], [0.3 + q, 0.7, -1.1]))
    for a in range(n):
        for b in range(n):
            if a == b:
                continue
            for k in two:
                cases.append((k, [a, b], []))
            for k in two_rot:
                cases.append((k, [a, b], [0.9 - 0.2 * a + 0.1 * b]))
    cases.append((abi.TOFFOLI, [0, 2, 4], []))
    cases.append((abi.TOFFOLI, [4, 1, 3], []))
    return cases


def main():
    out = {}
    # QFT, 12 qubits
    n = 12
    v = algorithms.normal_state_vector(n, 1)
    v = v / np.linalg.norm(v)
    gates = oracle.qft_gates(n)
    w = oracle.execute(v, n, gates)
    ref = np.sqrt(len(v)) * np.fft.ifft(v)
    assert np.max(np.abs(w - ref)) < 1e-12, "QFT disagrees with sqrt(N) ifft"
    out["qft12_in"], out["qft12_out"] = v, w
    # every gate on every target, 5 qubits
    n = 5
    rng = np.random.default_rng(7)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    v /= np.linalg.norm(v)
    desc, outs = [], []
    for kind, qs, th in gate_cases(n):
        g = abi.make_gate(kind, qs, th)
        w = oracle.apply_gate(v, n, g)
        want = bf.gate_matrix(n, g) @ v
        assert np.max(np.abs(w - want)) < 1e-13, (kind, qs)
        q = list(qs) + [-1] * (3 - len(qs))
        t = list(th) + [0.0] * (3 - len(th))
        desc.append([kind] + q + t)
        outs.append(w)
    out["gates_n"] = np.array(n)
    out["gates_in"] = v
    out["gates_desc"] = np.array(desc, dtype=np.float64)
    out["gates_out"] = np.array(outs)
    # Pauli sums with Y terms, both conventions
    n = 6
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    codes = rng.integers(0, 4, size=(20, 8, n)).astype(np.uint8)
    coeffs = rng.normal(size=(20, 8))
    eref, ephys = [], []
    for s in range(20):
        e = oracle.pauli_sum(psi, n, codes[s], coeffs[s])
        assert abs(e - bf.expectation_reference(psi, n, codes[s], coeffs[s])) < 1e-13
        eref.append(e)
        ephys.append(bf.expectation_physical(psi, n, codes[s], coeffs[s]))
    out["pauli_state"], out["pauli_codes"], out["pauli_coeffs"] = psi, codes, coeffs
    out["pauli_ref"], out["pauli_phys"] = np.array(eref), np.array(ephys)
    # parameter shift, 8 qubits
    n, layers = 8, 2
    tmpl, P = algorithms.ry_rz_ring_template(n, layers)
    x = rng.uniform(0, 2 * math.pi, P)
    pcodes, pcoeffs = algorithms.heisenberg_codes(n, 0.4, -0.9, 0.6)
    e = oracle.energy(n, tmpl, x, pcodes, pcoeffs)
    g = oracle.param_shift(n, tmpl, x, pcodes, pcoeffs)
    fd = oracle.finite_difference(n, tmpl, x, pcodes, pcoeffs, 1e-6)
    assert np.max(np.abs(g - fd)) < 1e-4, "parameter shift disagrees with finite differences"
    out["ps_params"], out["ps_energy"], out["ps_grad"] = x, np.array(e), g
    out["ps_codes"], out["ps_coeffs"] = pcodes, pcoeffs
    np.savez_compressed(os.path.join(HERE, "fixtures.npz"), **out)
    print("wrote", os.path.join(HERE, "fixtures.npz"))


if __name__ == "__main__":
    main()
