"""The committed golden fixtures (tests/golden/fixtures.npz, written by
tests/golden/make_golden.py after cross-checks against numpy brute force and
the QFT closed form) still match the oracle and the closed forms."""
import os

import numpy as np

import oracle
from logosq_amd import abi

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.npz")


def _fx():
    return np.load(FIX, allow_pickle=False)


def _gate(row):
    kind = int(row[0])
    qs = [int(q) for q in row[1:4] if q >= 0]
    return abi.make_gate(kind, qs, list(row[4:4 + 3]))


def test_qft_fixture():
    f = _fx()
    v, w = f["qft12_in"], f["qft12_out"]
    assert np.max(np.abs(w - np.sqrt(len(v)) * np.fft.ifft(v))) < 1e-12
    assert np.max(np.abs(oracle.execute(v, 12, oracle.qft_gates(12)) - w)) < 1e-15


def test_gate_fixtures_reproduce():
    f = _fx()
    n, v = int(f["gates_n"]), f["gates_in"]
    assert len(f["gates_desc"]) > 250
    for row, w in zip(f["gates_desc"], f["gates_out"]):
        assert np.max(np.abs(oracle.apply_gate(v, n, _gate(row)) - w)) < 1e-15


def test_pauli_and_gradient_fixtures_reproduce():
    f = _fx()
    psi = f["pauli_state"]
    for codes, coeffs, e in zip(f["pauli_codes"], f["pauli_coeffs"], f["pauli_ref"]):
        assert abs(oracle.pauli_sum(psi, 6, codes, coeffs) - e) < 1e-14
    # the two conventions differ exactly on terms with an odd number of Y
    assert np.max(np.abs(f["pauli_ref"] - f["pauli_phys"])) > 1e-3
    from logosq_amd import algorithms
    tmpl, P = algorithms.ry_rz_ring_template(8, 2)
    g = oracle.param_shift(8, tmpl, f["ps_params"], f["ps_codes"], f["ps_coeffs"])
    assert np.max(np.abs(g - f["ps_grad"])) < 1e-14
