"""CPU restatement of measurement (src/states/mod.rs:180-262,
src/circuits/mod.rs:215-233) with the caller-supplied randomness of the GPU
entry points: uniforms for measure_qubit, a SplitMix64 stream for shots.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker.
"""
import numpy as np

import oracle

MASK64 = (1 << 64) - 1


def splitmix64(seed, count):
    """SplitMix64 outputs (Steele, Lea, Flood 2014)."""
    s = seed & MASK64
    out = []
    for _ in range(count):
        s = (s + 0x9E3779B97F4A7C15) & MASK64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        out.append(z ^ (z >> 31))
    return out


def bit_probabilities(v, n):
    p = np.abs(np.asarray(v)) ** 2
    idx = np.arange(len(p), dtype=np.uint64)
    return np.array([p[(idx >> np.uint64(b)) & np.uint64(1) == 1].sum() for b in range(n)])


def measure_qubit(v, n, qubit_idx, u):
    """mod.rs:196-231: mask = 1 << qubit_idx; outcome 1 when u < P(1); keep
    the matching amplitudes; State::normalize."""
    v = np.array(v, dtype=np.complex128, copy=True)
    mask = 1 << qubit_idx
    idx = np.arange(len(v))
    p1 = float(np.sum(np.abs(v[(idx & mask) != 0]) ** 2))
    r = 1 if u < p1 else 0
    v[(idx & mask) != (mask if r else 0)] = 0
    v, _ = oracle.normalize(v)
    return r, v


def sample(v, seed, shots):
    """WeightedIndex rule: cumulative weights c (in order), x = u * c[-1],
    index = first i with c_i > x; u from SplitMix64(seed), top 53 bits."""
    c = np.cumsum(np.abs(np.asarray(v)) ** 2)
    us = np.array([(z >> 11) * (1.0 / 9007199254740992.0) for z in splitmix64(seed, shots)])
    x = us * c[-1]
    return np.minimum(np.searchsorted(c, x, side="right"), len(c) - 1)
