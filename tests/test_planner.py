"""The fused engine's planner, checked on the CPU: the order in which it
applies the gates (lq_plan_fused) must give the in-order result, which the
oracle replays; every gate is applied exactly once; renormalisations are
barriers.  No GPU needed (planning is host code)."""
import ctypes
import math

import numpy as np
import pytest

import oracle
import logosq_amd as lq
from logosq_amd import abi
from tests import bruteforce as bf
from tests.test_oracle import G


def plan(n, gates, tile_bits=0, low_bits=0, dtype=abi.LQ_C128):
    lib = lq.load()
    arr = abi.gate_array(gates)
    order = (ctypes.c_int * max(1, len(gates)))()
    pidx = (ctypes.c_int * max(1, len(gates)))()
    npass = ctypes.c_int()
    rc = lib.lq_plan_fused(n, dtype, arr, len(gates), tile_bits, low_bits, order, pidx, ctypes.byref(npass))
    assert rc == 0, lib.lq_last_error()
    return list(order)[:len(gates)], list(pidx)[:len(gates)], npass.value


def random_circuit(n, count, rng, with_dense=True):
    kinds1 = [abi.X, abi.Y, abi.Z, abi.H, abi.S, abi.T, abi.RX, abi.RY, abi.RZ, abi.PHASE, abi.U3]
    kinds2 = [abi.CNOT, abi.CZ, abi.SWAP, abi.CPHASE, abi.CRX, abi.CRY, abi.CRZ, abi.CP, abi.RXX, abi.RYY,
              abi.RZZ]
    out = []
    for _ in range(count):
        r = rng.uniform()
        if r < 0.45:
            k = kinds1[rng.integers(len(kinds1))]
            out.append(G(k, [int(rng.integers(n))], rng.uniform(0, 6.28, 3)))
        elif r < 0.9 and n >= 2:
            a, b = rng.choice(n, 2, replace=False)
            out.append(G(kinds2[rng.integers(len(kinds2))], [int(a), int(b)], [rng.uniform(0, 6.28)]))
        elif n >= 3:
            a, b, c = rng.choice(n, 3, replace=False)
            out.append(G(abi.TOFFOLI, [int(a), int(b), int(c)]))
        if with_dense and rng.uniform() < 0.03:
            m = bf.u3(*rng.uniform(0, 6, 3))
            out.append(G(abi.MATRIX1, [int(rng.integers(n))], (), m.reshape(-1), abi.GATE_FLAG_RENORMALIZE))
    return out


@pytest.mark.parametrize("n,tile,low", [(3, 0, 0), (6, 4, 2), (8, 5, 2), (9, 6, 3), (10, 6, 2), (12, 7, 3)])
def test_planned_order_equals_in_order(n, tile, low):
    rng = np.random.default_rng(n * 100 + tile)
    gates = random_circuit(n, 120, rng)
    order, pidx, npass = plan(n, gates, tile, low)
    assert sorted(order) == list(range(len(gates)))
    assert all(pidx[i] <= pidx[i + 1] for i in range(len(pidx) - 1))
    psi = bf.random_state(n, rng)
    want = oracle.execute(psi, n, gates)
    got = oracle.execute(psi, n, [gates[i] for i in order])
    assert np.max(np.abs(got - want)) < 1e-12
    # a renormalisation is a barrier: every gate up to a renormalising gate
    # runs in passes before every gate after it
    dense = [i for i, g in enumerate(gates) if g.flags & abi.GATE_FLAG_RENORMALIZE]
    pass_of = {g: pidx[k] for k, g in enumerate(order)}
    for d in dense:
        if d + 1 < len(gates):
            assert max(pass_of[i] for i in range(d + 1)) < min(pass_of[i] for i in range(d + 1, len(gates)))


def test_brickwork_pass_counts():
    """The 30-qubit depth-20 brickwork (890 gates) collapses to far fewer
    passes; each pass is one HBM read + write of the state."""
    circ = lq.algorithms.brickwork_circuit(30, 20, 2)
    d = circ.descriptors()
    order, pidx, npass = plan(30, d)
    assert sorted(order) == list(range(len(d)))
    print("brickwork 30q: %d gates -> %d passes" % (len(d), npass))
    assert npass < len(d) / 5


def test_small_state_single_pass():
    """A state that fits one tile runs its whole circuit in one pass."""
    circ = lq.algorithms.brickwork_circuit(10, 20, 5)
    order, pidx, npass = plan(10, circ.descriptors())
    assert npass == 1


@pytest.mark.parametrize("tile,low", [(12, 4), (12, 5), (13, 4)])
@pytest.mark.parametrize("n,seed", [(14, 1), (15, 2), (16, 3)])
def test_window_search_geometries(n, seed, tile, low):
    """The tile geometries the autotuner chooses between at 30 qubits, on
    states a few bits wider than a tile (so a pass's high bits are picked
    among several windows): a brickwork circuit and a random circuit, each
    planned order replayed by the oracle against the in-order result, every
    gate once, passes non-decreasing."""
    rng = np.random.default_rng(1000 * n + seed)
    brick = lq.algorithms.brickwork_circuit(n, 12, seed).descriptors()
    for gates in (brick, random_circuit(n, 160, rng, with_dense=False)):
        order, pidx, npass = plan(n, gates, tile, low)
        assert sorted(order) == list(range(len(gates)))
        assert all(pidx[i] <= pidx[i + 1] for i in range(len(pidx) - 1))
        assert 1 <= npass <= len(gates)
        psi = bf.random_state(n, rng)
        want = oracle.execute(psi, n, gates)
        got = oracle.execute(psi, n, [gates[i] for i in order])
        assert np.max(np.abs(got - want)) < 1e-12
