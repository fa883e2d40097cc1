"""Noise channels on the GPU (src/noise/mod.rs:10-380, Circuit::execute with
noise, src/circuits/mod.rs:352-375) against the noise oracle: amplitudes
<= 1e-12 (complex128), <= 1e-5 (complex64); depolarizing draws replayed
from the same SplitMix64 stream."""
import numpy as np
import pytest

import logosq_amd as lq
import oracle
from logosq_amd import algorithms
from oracle import noise as on

pytestmark = pytest.mark.gpu
N = 7


def _state(seed, dtype=lq.C128):
    v = algorithms.normal_state_vector(N, seed)
    return lq.State.new(v, dtype=dtype), oracle.normalize(v)[0]


@pytest.mark.parametrize("dtype,tol", [(lq.C128, 1e-12), (lq.C64, 1e-5)])
def test_damping_channels(hip, dtype, tol):
    for targets in (None, [0], [6, 2, 4], [3, 3]):
        print("noise: damping targets", targets, "dtype", dtype, flush=True)
        st, ref = _state(31, dtype)
        lq.AmplitudeDampingNoise(0.3).with_target_qubits(targets if targets is not None else range(N)).apply(st)
        assert np.max(np.abs(st.vector() - on.amplitude_damping(ref, N, 0.3, targets))) <= tol
        st, ref = _state(32, dtype)
        m = lq.PhaseDampingNoise(0.45)
        if targets is not None:
            m.with_target_qubits(targets)
        m.apply(st)
        assert np.max(np.abs(st.vector() - on.phase_damping(ref, N, 0.45, targets))) <= tol
    st, ref = _state(33, dtype)
    lq.ThermalRelaxationNoise(40.0, 25.0, 3.0).with_target_qubits([1, 5]).apply(st)
    assert np.max(np.abs(st.vector() - on.thermal_relaxation(ref, N, 40.0, 25.0, 3.0, [1, 5]))) <= tol
    st, _ = _state(34, dtype)
    lq.AmplitudeDampingNoise(0.3).with_target_qubits([]).apply(st)
    assert np.all(st.vector() == 0)


def test_depolarizing(hip):
    for p, seed in ((1.0, 5), (0.35, 6), (0.0, 7)):
        print("noise: depolarizing p", p, flush=True)
        st, ref = _state(40)
        m = lq.DepolarizingNoise(p, seed=seed)
        m.apply(st)
        want, applied = on.depolarizing(ref, N, p, on.SplitMix64(seed))
        assert m.errors_applied == applied
        assert np.max(np.abs(st.vector() - want)) <= 1e-12


def test_circuit_with_noise(hip):
    """Circuit::execute with a composite model after every operation."""
    print("noise: circuit", flush=True)
    c = lq.Circuit(N)
    for q in range(N):
        c.h(q)
        c.rx(q, 0.2 + 0.1 * q)
    for q in range(N - 1):
        c.cnot(q, q + 1)
    comp = lq.CompositeNoise()
    comp.add_noise(lq.AmplitudeDampingNoise(0.05).with_target_qubits([0, 3]))
    comp.add_noise(lq.DepolarizingNoise(0.1, seed=99))
    comp.add_noise(lq.PhaseDampingNoise(0.02))
    c.add_noise(comp)
    st = lq.State.zero_state(N)
    c.execute(st)
    rng = on.SplitMix64(99)
    models = [lambda v: on.amplitude_damping(v, N, 0.05, [0, 3]),
              lambda v: on.depolarizing(v, N, 0.1, rng)[0],
              lambda v: on.phase_damping(v, N, 0.02)]
    want = on.execute_with_noise(oracle.zero_state(N), N, c.descriptors(), models)
    assert np.max(np.abs(st.vector() - want)) <= 1e-12
    # without noise models execute equals execute_without_noise
    st2 = lq.State.zero_state(N)
    c.execute_without_noise(st2)
    assert np.max(np.abs(st2.vector() - oracle.execute(oracle.zero_state(N), N, c.descriptors()))) <= 1e-12


def test_noise_argument_errors(hip):
    st = lq.State.zero_state(3)
    with pytest.raises(lq.InvalidQubitIndex):
        lq.AmplitudeDampingNoise(0.1).with_target_qubits([3]).apply(st)
    with pytest.raises(lq.LogosQError):
        lq.PhaseDampingNoise(1.5).apply(st)
