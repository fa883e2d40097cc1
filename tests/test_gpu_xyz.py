"""XYZ Heisenberg Trotter evolution on the GPU (src/algorithms/xyz_heisenberg.rs
:31-162, examples/xyz_h.rs): amplitudes <= 1e-12 and energies <= 1e-10
against the oracle, per-gate and fused; the reference's known answers."""
import numpy as np
import pytest

import logosq_amd as lq
import oracle
from logosq_amd import abi, algorithms, xyz_h

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [abi.EXEC_PER_GATE, abi.EXEC_FUSED])
def test_trotter_parity(hip, mode):
    n = 9
    print("xyz: trotter n=%d mode %d" % (n, mode), flush=True)
    psi0 = algorithms.normal_state_vector(n, 77)
    p = algorithms.HeisenbergParameters(0.9, -0.35, 0.6, 0.25, time_steps=4, dt=0.12)
    st = lq.State.new(psi0)
    algorithms.xyz_heisenberg_circuit(n, p).execute(st, mode)
    ref0, _ = oracle.normalize(psi0)
    ref = oracle.execute(ref0, n, algorithms.xyz_heisenberg_circuit(n, p).descriptors())
    assert np.max(np.abs(st.vector() - ref)) <= 1e-12
    e = algorithms.xyz_energy_efficient(st, p)
    assert abs(e - oracle.xyz_energy(ref, n, p.jx, p.jy, p.jz, p.external_field)) <= 1e-10


def test_reference_known_answers(hip):
    print("xyz: known answers", flush=True)
    # test_mps_field_only_evolution_matches_dense_result (tests/test_simulators_mps.rs:155-196)
    p = algorithms.HeisenbergParameters(0.0, 0.0, 0.0, 0.35, time_steps=4, dt=0.2)
    st = lq.State.one_state(3)
    algorithms.xyz_heisenberg_simulate(st, p)
    v = st.vector()
    assert abs(v[7] - np.exp(1j * 3 * 4 * 0.35 * 0.2 / 2)) <= 1e-12
    assert abs(algorithms.xyz_energy_efficient(st, p) + 1.05) <= 1e-10
    # test_mps_initial_state_expectation_matches_field_energy (tests/algorithms/test_xyz_heisenberg.rs:6-27)
    p = algorithms.HeisenbergParameters(0.0, 0.0, 0.0, 1.0, time_steps=0, dt=0.1)
    assert abs(algorithms.xyz_energy_efficient(lq.State.one_state(4), p) + 4.0) <= 1e-10


def test_xyz_h_harness(hip):
    """The dense record of examples/xyz_h.rs (defaults: 8 qubits, 10 steps,
    dt 0.1, J = 1, no field, from |1...1>) against the oracle."""
    print("xyz: harness", flush=True)
    rec = xyz_h.run(num_qubits=8, time_steps=10, dt=0.1, jx=1.0, jy=1.0, jz=1.0, external_field=0.2)
    p = algorithms.HeisenbergParameters(1.0, 1.0, 1.0, 0.2, time_steps=10, dt=0.1)
    psi = np.zeros(256, dtype=complex)
    psi[255] = 1
    ref = oracle.execute(psi, 8, algorithms.xyz_heisenberg_circuit(8, p).descriptors())
    assert abs(rec["initial_energy"] - oracle.xyz_energy(psi, 8, 1.0, 1.0, 1.0, 0.2)) <= 1e-10
    assert abs(rec["final_energy"] - oracle.xyz_energy(ref, 8, 1.0, 1.0, 1.0, 0.2)) <= 1e-10
    assert rec["num_operations"] == 10 * (3 * 7 + 8)
