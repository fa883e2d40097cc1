"""Noise models: the mirror of src/noise/mod.rs (NoiseModel, NoiseParams,
DepolarizingNoise, AmplitudeDampingNoise, PhaseDampingNoise,
ThermalRelaxationNoise, CompositeNoise), applied on the device state by
lq_noise_* in the C ABI.

Qubit numbers address bit `1 << qubit` of the basis index, as the
reference's noise loops do (mod.rs:129,197,343-376), not bit n-1-qubit as
the gates do.  The reference's depolarizing noise draws from thread_rng; here
each DepolarizingNoise owns a SplitMix64 state (seed given at construction),
so runs are reproducible.
"""
import ctypes
import math

import numpy as np

from ._lib import call


def _qubit_array(state, targets):
    qs = list(range(state.num_qubits())) if targets is None else [int(q) for q in targets]
    return (ctypes.c_int * max(len(qs), 1))(*qs), len(qs)


class NoiseModel:
    """trait NoiseModel (mod.rs:10-17)."""

    def apply(self, state):
        raise NotImplementedError

    def name(self):
        raise NotImplementedError


class NoiseParams:
    """NoiseParams (mod.rs:19-40): error probability and target qubits
    (None: every qubit)."""

    def __init__(self, error_probability):
        self.error_probability = float(error_probability)
        self.target_qubits = None

    def with_target_qubits(self, qubits):
        self.target_qubits = [int(q) for q in qubits]
        return self


class DepolarizingNoise(NoiseModel):
    """DepolarizingNoise (mod.rs:42-91): per target qubit, with probability p
    one of X, Y (the reference's i / -i form), Z."""

    def __init__(self, error_probability, seed=0):
        self.params = NoiseParams(error_probability)
        self.rng_state = ctypes.c_uint64(int(seed) & ((1 << 64) - 1))
        self.errors_applied = 0

    def with_target_qubits(self, qubits):
        self.params.with_target_qubits(qubits)
        return self

    def apply(self, state):
        arr, count = _qubit_array(state, self.params.target_qubits)
        applied = ctypes.c_int()
        call("lq_noise_depolarizing", state.handle, arr, count, self.params.error_probability,
             ctypes.byref(self.rng_state), ctypes.byref(applied))
        self.errors_applied += applied.value

    def name(self):
        return "Depolarizing(p=%s)" % _fmt(self.params.error_probability)


class AmplitudeDampingNoise(NoiseModel):
    """AmplitudeDampingNoise (mod.rs:93-152)."""

    def __init__(self, error_probability):
        self.params = NoiseParams(error_probability)

    def with_target_qubits(self, qubits):
        self.params.with_target_qubits(qubits)
        return self

    def apply(self, state):
        arr, count = _qubit_array(state, self.params.target_qubits)
        call("lq_noise_amplitude_damping", state.handle, arr, count, self.params.error_probability)

    def name(self):
        return "AmplitudeDamping(p=%s)" % _fmt(self.params.error_probability)


class PhaseDampingNoise(NoiseModel):
    """PhaseDampingNoise (mod.rs:154-211)."""

    def __init__(self, error_probability):
        self.params = NoiseParams(error_probability)

    def with_target_qubits(self, qubits):
        self.params.with_target_qubits(qubits)
        return self

    def apply(self, state):
        arr, count = _qubit_array(state, self.params.target_qubits)
        call("lq_noise_phase_damping", state.handle, arr, count, self.params.error_probability)

    def name(self):
        return "PhaseDamping(p=%s)" % _fmt(self.params.error_probability)


class ThermalRelaxationNoise(NoiseModel):
    """ThermalRelaxationNoise (mod.rs:213-300): amplitude damping with
    p_a = 1 - exp(-t/T1), then phase damping with p_p = 1 - exp(-t/T_phi),
    1/T_phi = 1/T2 - 1/(2 T1); T2 is capped at 2 T1 on construction."""

    def __init__(self, t1, t2, gate_time):
        self.t1 = float(t1)
        self.t2 = min(float(t2), 2.0 * float(t1))
        self.gate_time = float(gate_time)
        self.target_qubits = None

    def with_target_qubits(self, qubits):
        self.target_qubits = [int(q) for q in qubits]
        return self

    def probabilities(self):
        p_amp = 1.0 - math.exp(-self.gate_time / self.t1)
        rate = 1.0 / self.t2 - 1.0 / (2.0 * self.t1) if self.t2 <= 2.0 * self.t1 else 0.0
        p_phase = 1.0 - math.exp(-self.gate_time * rate) if rate > 0.0 else 0.0
        return p_amp, p_phase

    def apply(self, state):
        p_amp, p_phase = self.probabilities()
        targets = list(range(state.num_qubits())) if self.target_qubits is None else self.target_qubits
        if p_amp > 0.0:
            AmplitudeDampingNoise(p_amp).with_target_qubits(targets).apply(state)
        if p_phase > 0.0:
            PhaseDampingNoise(p_phase).with_target_qubits(targets).apply(state)

    def name(self):
        return "ThermalRelaxation(T1=%s, T2=%s, gate_time=%s)" % (_fmt(self.t1), _fmt(self.t2), _fmt(self.gate_time))


class CompositeNoise(NoiseModel):
    """CompositeNoise (mod.rs:302-336): the models in sequence."""

    def __init__(self):
        self.noise_models = []

    def add_noise(self, model):
        self.noise_models.append(model)
        return self

    def apply(self, state):
        for m in self.noise_models:
            m.apply(state)

    def name(self):
        return "Composite(%s)" % ", ".join(m.name() for m in self.noise_models)


def _fmt(x):
    """Rust's Display for f64: shortest round-trip form, no trailing '.0'
    for integral values."""
    x = float(x)
    if np.isfinite(x) and x == int(x) and abs(x) < 1e16:
        return str(int(x))
    return repr(x)
