"""Pins the measurement oracle: SplitMix64 known answers, bit marginals and
measure_qubit against brute force (including the reference's bit-order:
mask = 1 << qubit_idx), and the sampling rule against exact frequencies."""
import numpy as np

from oracle import measure as om


def test_splitmix64_known_answers():
    # published first outputs of SplitMix64 for seed 0 and seed 1234567
    assert om.splitmix64(0, 1)[0] == 0xE220A8397B1DCDAF
    assert om.splitmix64(1234567, 1)[0] == 6457827717110365317


def test_bit_probabilities_and_measure_qubit_brute_force():
    rng = np.random.default_rng(2)
    n = 4
    v = rng.normal(size=16) + 1j * rng.normal(size=16)
    v /= np.linalg.norm(v)
    p1 = om.bit_probabilities(v, n)
    for b in range(n):
        assert abs(p1[b] - sum(abs(v[i]) ** 2 for i in range(16) if (i >> b) & 1)) < 1e-15
    # qubit_idx 0 is bit 0 (the least significant), not qubit 0 of the gates
    r, w = om.measure_qubit(v, n, 0, 0.0 if p1[0] > 0 else 1.0)
    assert r == 1 and all(abs(w[i]) == 0 for i in range(16) if i % 2 == 0)
    assert abs(np.linalg.norm(w) - 1.0) < 1e-15
    r, w = om.measure_qubit(v, n, 3, 0.999999)
    assert r == 0 and all(abs(w[i]) == 0 for i in range(8, 16))


def test_sampling_rule_frequencies():
    v = np.sqrt(np.array([0.1, 0.0, 0.6, 0.3]))
    s = om.sample(v, 7, 20000)
    f = np.bincount(s, minlength=4) / len(s)
    assert f[1] == 0.0
    assert np.max(np.abs(f - [0.1, 0.0, 0.6, 0.3])) < 0.015
