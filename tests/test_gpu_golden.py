"""The GPU path against the committed golden fixtures: amplitudes <= 1e-12,
expectations (both Y conventions) and gradients <= 1e-10."""
import os

import numpy as np
import pytest

import logosq_amd as lq
from logosq_amd import abi, algorithms

pytestmark = pytest.mark.gpu
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.npz")


def test_golden_fixtures_on_gpu(hip):
    f = np.load(FIX, allow_pickle=False)
    print("golden: qft", flush=True)
    st = lq.State.new(f["qft12_in"])
    algorithms.qft_circuit(12).execute(st)
    assert np.max(np.abs(st.vector() - f["qft12_out"])) <= 1e-12
    print("golden: gates", flush=True)
    n, v = int(f["gates_n"]), f["gates_in"]
    for row, w in zip(f["gates_desc"], f["gates_out"]):
        g = abi.make_gate(int(row[0]), [int(q) for q in row[1:4] if q >= 0], list(row[4:7]))
        st = lq.State.new(v)
        lq._lib.call("lq_apply_gate", st.handle, __import__("ctypes").byref(g))
        assert np.max(np.abs(st.vector() - w)) <= 1e-12, row
    print("golden: pauli", flush=True)
    st = lq.State.new(f["pauli_state"])
    for conv, key in ((abi.PAULI_REFERENCE, "pauli_ref"), (abi.PAULI_PHYSICAL, "pauli_phys")):
        for codes, coeffs, e in zip(f["pauli_codes"], f["pauli_coeffs"], f[key]):
            obs = lq.PauliObservable(6, conv)
            for row, c in zip(codes, coeffs):
                obs.add_term(lq.PauliTerm(float(c), [int(p) for p in row]))
            assert abs(obs.expectation(st) - e) <= 1e-10
    print("golden: gradient", flush=True)
    ans = lq.ry_rz_cz_ring_ansatz(8, 2)
    obs = lq.heisenberg_ring(8, 0.4, -0.9, 0.6)
    for ps in (lq.ParameterShift(), lq.ParameterShift(batched=True)):
        e, g = ps.compute_gradient(ans, obs, f["ps_params"], energy=True)
        assert abs(e - float(f["ps_energy"])) <= 1e-10
        assert np.max(np.abs(np.array(g) - f["ps_grad"])) <= 1e-10
