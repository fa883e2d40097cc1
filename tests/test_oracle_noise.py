"""Pins the noise oracle (oracle/noise.py, restating src/noise/mod.rs:10-380):
Kraus-operator brute force built with Kronecker products, closed forms and
the reference's quirks (sum of Kraus operators, last-target phase damping,
low-order bit numbering).  CPU only."""
import math
from functools import reduce

import numpy as np
import pytest

from oracle import noise as on
from tests import bruteforce as bf

I2 = np.eye(2, dtype=complex)


def on_bit(n, q, m):
    """m on bit q of the basis index (bit 0 = rightmost Kronecker factor)."""
    return reduce(np.kron, [m if (n - 1 - k) == q else I2 for k in range(n)])


def normalized(v):
    nv = np.linalg.norm(v)
    return v / nv if nv > 1e-10 else v


@pytest.mark.parametrize("targets", [None, [0], [2, 0, 3], [1, 1]])
def test_amplitude_damping_kraus_sum(targets):
    n, g = 4, 0.3
    rng = np.random.default_rng(11)
    psi = bf.random_state(n, rng)
    k0 = np.array([[1, 0], [0, math.sqrt(1 - g)]])
    k1 = np.array([[0, math.sqrt(g)], [0, 0]])
    qs = range(n) if targets is None else targets
    want = normalized(sum(on_bit(n, q, k0 + k1) @ psi for q in qs))
    assert np.max(np.abs(on.amplitude_damping(psi, n, g, targets) - want)) < 1e-13


def test_amplitude_damping_one_qubit_closed_form():
    a, b, g = 0.6, 0.8j, 0.25
    got = on.amplitude_damping([a, b], 1, g)
    want = normalized(np.array([a + math.sqrt(g) * b, math.sqrt(1 - g) * b]))
    assert np.max(np.abs(got - want)) < 1e-15


def test_phase_damping_last_target_only():
    n, lam = 4, 0.4
    psi = bf.random_state(n, np.random.default_rng(12))
    d = np.diag([1, math.sqrt(1 - lam)])
    got = on.phase_damping(psi, n, lam, [3, 1])
    assert np.max(np.abs(got - normalized(on_bit(n, 1, d) @ psi))) < 1e-13
    assert np.max(np.abs(got - normalized(on_bit(n, 3, d) @ psi))) > 1e-3


def test_bit_order_is_low_order():
    """A planted mistake: numbering bits from the top (n-1-q) must differ."""
    n, g = 3, 0.5
    psi = bf.random_state(n, np.random.default_rng(13))
    k = np.array([[1, math.sqrt(g)], [0, math.sqrt(1 - g)]])
    msb = normalized(on_bit(n, n - 1 - 0, k) @ psi)
    assert np.max(np.abs(on.amplitude_damping(psi, n, g, [0]) - msb)) > 1e-3


def test_empty_target_list_gives_zero_vector():
    psi = bf.random_state(3, np.random.default_rng(14))
    assert np.all(on.amplitude_damping(psi, 3, 0.2, []) == 0)
    assert np.all(on.phase_damping(psi, 3, 0.2, []) == 0)


def test_depolarizing_replays_draws():
    n = 4
    psi = bf.random_state(n, np.random.default_rng(15))
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    negY = np.array([[0, 1j], [-1j, 0]])   # apply_y_error's form: -Y
    Z = np.diag([1.0 + 0j, -1.0])
    for p, seed in ((1.0, 7), (0.5, 8), (0.0, 9)):
        got, applied = on.depolarizing(psi, n, p, on.SplitMix64(seed))
        r, want, cnt = on.SplitMix64(seed), psi.copy(), 0
        for q in range(n):
            if r.uniform() < p:
                e = min(2, int(3 * r.uniform()))
                want = on_bit(n, q, (X, negY, Z)[e]) @ want
                cnt += 1
        assert applied == cnt and np.max(np.abs(got - want)) < 1e-15
        if p == 1.0:
            assert applied == n
        if p == 0.0:
            assert applied == 0 and np.array_equal(got, psi)


def test_depolarizing_rate():
    r = on.SplitMix64(3)
    hits = sum(on.depolarizing(np.array([1, 0], dtype=complex), 1, 0.2, r)[1] for _ in range(5000))
    assert abs(hits / 5000 - 0.2) < 0.02


def test_thermal_probabilities_closed_form():
    pa, pp = on.thermal_probabilities(50.0, 30.0, 1.0)
    assert abs(pa - (1 - math.exp(-1 / 50))) < 1e-15
    assert abs(pp - (1 - math.exp(-(1 / 30 - 1 / 100)))) < 1e-15
    # T2 above 2 T1 is capped to 2 T1: no pure dephasing left
    assert on.thermal_probabilities(10.0, 50.0, 1.0)[1] == 0.0
