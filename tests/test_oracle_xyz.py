"""XYZ Heisenberg Trotter evolution (src/algorithms/xyz_heisenberg.rs:31-162)
on the CPU oracle, pinned by Kronecker brute force, the reference's own
known answers and closed forms.  CPU only."""
import math

import numpy as np

import oracle
from logosq_amd import abi, algorithms
from tests import bruteforce as bf


def _run_oracle(n, params, psi):
    return oracle.execute(psi, n, algorithms.xyz_heisenberg_circuit(n, params).descriptors())


def test_builder_gate_list():
    """create_circuit (xyz_heisenberg.rs:31-61): per step the XX bonds, then
    YY, then ZZ (angle 2*J*dt), then RZ(h*dt) on every qubit."""
    p = algorithms.HeisenbergParameters(0.3, -0.2, 0.7, 0.5, time_steps=2, dt=0.1)
    d = algorithms.xyz_heisenberg_circuit(4, p).descriptors()
    assert len(d) == 2 * (3 * 3 + 4)
    step = [(g.kind, tuple(g.qubits)[:2], g.theta[0]) for g in d[:13]]
    want = [(abi.RXX, (i, i + 1), 2 * 0.3 * 0.1) for i in range(3)] + \
        [(abi.RYY, (i, i + 1), 2 * -0.2 * 0.1) for i in range(3)] + \
        [(abi.RZZ, (i, i + 1), 2 * 0.7 * 0.1) for i in range(3)]
    for (k, q, t), (wk, wq, wt) in zip(step[:9], want):
        assert k == wk and q == wq and abs(t - wt) < 1e-15
    assert all(k == abi.RZ and abs(t - 0.05) < 1e-15 for k, _, t in step[9:])
    # no field: no RZ
    p0 = algorithms.HeisenbergParameters(1, 1, 1, 0.0, time_steps=3, dt=0.1)
    assert len(algorithms.xyz_heisenberg_circuit(5, p0).descriptors()) == 3 * 12


def test_trotter_matches_kron_brute_force():
    n = 4
    rng = np.random.default_rng(21)
    psi = bf.random_state(n, rng)
    p = algorithms.HeisenbergParameters(*rng.uniform(-1, 1, 3), external_field=0.4, time_steps=3, dt=0.15)
    want = psi.copy()
    for g in algorithms.xyz_heisenberg_circuit(n, p).descriptors():
        want = bf.gate_matrix(n, g) @ want
    got = _run_oracle(n, p, psi)
    assert np.max(np.abs(got - want)) < 1e-12


def test_field_only_closed_form():
    """test_mps_field_only_evolution_matches_dense_result
    (tests/test_simulators_mps.rs:155-196): J = 0, h = 0.35, 4 steps of
    dt = 0.2 on |111>.  RZ(t)|1> = e^{it/2}|1>, so the amplitude is
    e^{i*3*4*0.035} and the energy is -3h."""
    n = 3
    p = algorithms.HeisenbergParameters(0.0, 0.0, 0.0, 0.35, time_steps=4, dt=0.2)
    psi = np.zeros(8, dtype=complex)
    psi[7] = 1
    got = _run_oracle(n, p, psi)
    assert abs(got[7] - np.exp(1j * 3 * 4 * 0.35 * 0.2 / 2)) < 1e-12
    assert np.max(np.abs(got[:7])) < 1e-12
    assert abs(oracle.xyz_energy(got, n, 0.0, 0.0, 0.0, 0.35) + 1.05) < 1e-12


def test_one_state_field_energy():
    """test_mps_initial_state_expectation_matches_field_energy
    (tests/algorithms/test_xyz_heisenberg.rs:6-27): <1111| h sum Z |1111> = -4h."""
    psi = np.zeros(16, dtype=complex)
    psi[15] = 1
    assert abs(oracle.xyz_energy(psi, 4, 0.0, 0.0, 0.0, 1.0) + 4.0) < 1e-12


def test_orientation_rule_two_qubits():
    """With the reference's RXX/RYY orientation rule (qubit1 < qubit2 leaves
    the 01<->10 block alone, src/gates/parameterized.rs:563-575,638-650) a
    single bond's RXX(t) RYY(t) is the identity when Jx = Jy: on 00/11 the
    two blocks are inverse rotations.  The physical orientation would swap
    weight between 01 and 10."""
    rng = np.random.default_rng(5)
    psi = bf.random_state(2, rng)
    p = algorithms.HeisenbergParameters(0.6, 0.6, 0.0, 0.0, time_steps=5, dt=0.3)
    assert np.max(np.abs(_run_oracle(2, p, psi) - psi)) < 1e-12
    phys = bf.gate_matrix(2, abi.make_gate(abi.RXX, [1, 0], [0.36])) @ \
        bf.gate_matrix(2, abi.make_gate(abi.RYY, [1, 0], [0.36])) @ psi
    assert np.max(np.abs(phys - psi)) > 1e-3


def test_diagonal_evolution_conserves():
    """Jx = Jy = 0: every Trotter gate is diagonal, so |amplitudes| and the
    energy of the diagonal Hamiltonian are unchanged."""
    n = 6
    rng = np.random.default_rng(6)
    psi = bf.random_state(n, rng)
    p = algorithms.HeisenbergParameters(0.0, 0.0, -0.8, 0.3, time_steps=5, dt=0.25)
    got = _run_oracle(n, p, psi)
    assert np.max(np.abs(np.abs(got) - np.abs(psi))) < 1e-12
    e0 = oracle.xyz_energy(psi, n, 0.0, 0.0, -0.8, 0.3)
    assert abs(oracle.xyz_energy(got, n, 0.0, 0.0, -0.8, 0.3) - e0) < 1e-12
