"""CPU restatement of the noise models (src/noise/mod.rs:10-380) and of
Circuit::execute with noise (src/circuits/mod.rs:352-375), loop for loop,
with the depolarizing draws taken from a SplitMix64 stream (the GPU entry
point's randomness) instead of thread_rng.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker.
"""
import math

import numpy as np

import oracle

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK64

    def uniform(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        z ^= z >> 31
        return (z >> 11) * (1.0 / 9007199254740992.0)


def _targets(n, qubits):
    return list(range(n)) if qubits is None else list(qubits)


def amplitude_damping(v, n, gamma, qubits=None):
    """AmplitudeDampingNoise::apply (mod.rs:110-147): new accumulates
    (K0 + K1) psi for every target from the same psi, then normalize."""
    psi = np.asarray(v, dtype=np.complex128)
    new = np.zeros_like(psi)
    sg, s1 = math.sqrt(gamma), math.sqrt(1.0 - gamma)
    for q in _targets(n, qubits):
        m = 1 << q
        for i in range(len(psi)):
            if i & m == 0:
                new[i] += psi[i]
            else:
                new[i] += psi[i] * complex(s1, 0.0)
                new[i & ~m] += psi[i] * complex(sg, 0.0)
    return oracle.normalize(new)[0]


def phase_damping(v, n, lam, qubits=None):
    """PhaseDampingNoise::apply (mod.rs:173-206): the output is reset from
    psi for every target, so the last one decides; then normalize."""
    psi = np.asarray(v, dtype=np.complex128)
    new = np.zeros_like(psi)
    s = math.sqrt(1.0 - lam)
    for q in _targets(n, qubits):
        m = 1 << q
        for i in range(len(psi)):
            new[i] = psi[i]
            if i & m:
                new[i] *= complex(s, 0.0)
    return oracle.normalize(new)[0]


def pauli_error(v, q, kind):
    """apply_x_error / apply_y_error / apply_z_error (mod.rs:338-380)."""
    psi = np.asarray(v, dtype=np.complex128)
    out = psi.copy()
    m = 1 << q
    for i in range(len(psi)):
        if kind == 0:
            out[i] = psi[i ^ m]
        elif kind == 1:
            out[i] = (-1j if i & m else 1j) * psi[i ^ m]
        else:
            out[i] = -psi[i] if i & m else psi[i]
    return out


def depolarizing(v, n, p, rng, qubits=None):
    """DepolarizingNoise::apply (mod.rs:63-86) with uniforms from rng:
    u < p, then the error type floor(3u') capped at 2.  Returns (state,
    errors applied)."""
    psi = np.asarray(v, dtype=np.complex128)
    applied = 0
    for q in _targets(n, qubits):
        if rng.uniform() < p:
            e = min(2, int(rng.uniform() * 3.0))
            psi = pauli_error(psi, q, e)
            applied += 1
    return psi, applied


def thermal_probabilities(t1, t2, gate_time):
    """ThermalRelaxationNoise (mod.rs:232-270): T2 capped at 2 T1."""
    t2 = min(t2, 2.0 * t1)
    p_amp = 1.0 - math.exp(-gate_time / t1)
    rate = 1.0 / t2 - 1.0 / (2.0 * t1) if t2 <= 2.0 * t1 else 0.0
    p_phase = 1.0 - math.exp(-gate_time * rate) if rate > 0.0 else 0.0
    return p_amp, p_phase


def thermal_relaxation(v, n, t1, t2, gate_time, qubits=None):
    p_amp, p_phase = thermal_probabilities(t1, t2, gate_time)
    psi = np.asarray(v, dtype=np.complex128)
    if p_amp > 0.0:
        psi = amplitude_damping(psi, n, p_amp, qubits)
    if p_phase > 0.0:
        psi = phase_damping(psi, n, p_phase, qubits)
    return psi


def execute_with_noise(v, n, gates, models):
    """Circuit::execute with noise (src/circuits/mod.rs:352-375): each gate,
    then every model in order.  models: callables state -> state."""
    psi = np.asarray(v, dtype=np.complex128)
    for g in gates:
        psi = oracle.apply_gate(psi, n, g)
        for f in models:
            psi = f(psi)
    return psi
