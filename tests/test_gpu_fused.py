"""GPU parity of the fused tile engine (Circuit::execute in LDS tile passes)
against the oracle's in-order execution: complex128 <= 1e-12, complex64
<= 2e-6 (f32 rounding), on random circuits of every gate kind, several tile
geometries, and states from one tile to many tiles."""
import numpy as np
import pytest

import oracle
import logosq_amd as lq
from logosq_amd import abi
from tests import bruteforce as bf
from tests.test_planner import random_circuit

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def run(psi, n, gates, dtype, mode):
    st = lq.State(n, dtype)
    st.set_vector(psi)
    lq._lib.call("lq_execute", st.handle, abi.gate_array(gates), len(gates), mode)
    return st.vector()


@pytest.fixture
def tile_config():
    yield lambda a, b, c: lq._lib.call("lq_set_fused_config", a, b, c)
    lq._lib.call("lq_set_fused_config", 0, 0, 0)


@pytest.mark.parametrize("n", [1, 3, 8, 12, 13, 15, 18])
def test_fused_random_circuits(hip, n):
    say("fused n=%d" % n)
    rng = np.random.default_rng(500 + n)
    gates = random_circuit(n, 150, rng)
    psi = bf.random_state(n, rng)
    want = oracle.execute(psi, n, gates)
    got = run(psi, n, gates, lq.C128, abi.EXEC_FUSED)
    assert np.max(np.abs(got - want)) <= 1e-12
    got64 = run(psi, n, gates, lq.C64, abi.EXEC_FUSED)
    assert np.max(np.abs(got64 - want)) <= 2e-6


@pytest.mark.parametrize("tile,low", [(7, 2), (9, 4), (13, 6)])
def test_fused_tile_geometries(hip, tile_config, tile, low):
    say("fused geometry tile=%d low=%d" % (tile, low))
    tile_config(tile, tile, low)
    n = 16
    rng = np.random.default_rng(tile * 31 + low)
    gates = random_circuit(n, 120, rng, with_dense=False)
    psi = bf.random_state(n, rng)
    want = oracle.execute(psi, n, gates)
    assert np.max(np.abs(run(psi, n, gates, lq.C128, abi.EXEC_FUSED) - want)) <= 1e-12
    assert np.max(np.abs(run(psi, n, gates, lq.C64, abi.EXEC_FUSED) - want)) <= 2e-6


def test_fused_brickwork_22(hip):
    """The brickwork generator at 22 qubits, depth 20: fused vs per-gate on the
    GPU and vs the oracle."""
    say("fused brickwork 22")
    n = 22
    circ = lq.algorithms.brickwork_circuit(n, 20, 2)
    d = circ.descriptors()
    psi = oracle.zero_state(n)
    want = oracle.execute(psi, n, d)
    got = run(psi, n, d, lq.C128, abi.EXEC_FUSED)
    assert np.max(np.abs(got - want)) <= 1e-12
    per = run(psi, n, d, lq.C128, abi.EXEC_PER_GATE)
    assert np.max(np.abs(per - want)) <= 1e-12


def test_fused_brickwork_22_c64(hip):
    """The same circuit in complex64 (fused and per-gate) against the c128
    oracle: amplitudes <= 2e-6 (f32 storage), relative 2-norm error <= 1e-5."""
    say("fused brickwork 22 c64")
    n = 22
    d = lq.algorithms.brickwork_circuit(n, 20, 2).descriptors()
    psi = oracle.zero_state(n)
    want = oracle.execute(psi, n, d)
    for mode in (abi.EXEC_FUSED, abi.EXEC_PER_GATE):
        got = run(psi, n, d, lq.C64, mode)
        err = float(np.max(np.abs(got - want)))
        rel = float(np.linalg.norm(got - want))
        say("  c64 mode %d: max |delta| %.3e, relative 2-norm %.3e" % (mode, err, rel))
        assert err <= 2e-6
        assert rel <= 1e-5


def test_fused_qft_12(hip):
    say("fused qft12")
    n = 12
    psi0 = lq.algorithms.normal_state_vector(n, 1)
    st = lq.State.new(psi0)
    lq.algorithms.qft_circuit(n).execute(st, abi.EXEC_FUSED)
    ref = oracle.execute(oracle.normalize(psi0)[0], n, oracle.qft_gates(n))
    assert np.max(np.abs(st.vector() - ref)) <= 1e-12


@pytest.mark.parametrize("tile,low", [(12, 4), (12, 5), (13, 4)])
def test_fused_autotuner_geometries(hip, tile_config, tile, low):
    """The geometries the autotuner chooses between at 30 qubits, on a
    17-qubit state (a pass's high bits picked among several windows): a
    depth-12 brickwork on the run-time specialised pass kernels and a random
    circuit on the generic tile kernel, complex128 <= 1e-12 and complex64
    <= 2e-6 against the oracle."""
    from logosq_amd import jit
    from tests import modes
    say("fused autotuner geometry tile=%d low=%d" % (tile, low))
    tile_config(tile, tile, low)
    n = 17
    rng = np.random.default_rng(7000 + 10 * tile + low)
    brick = lq.algorithms.brickwork_circuit(n, 12, tile + low).descriptors()
    rand = random_circuit(n, 120, rng, with_dense=False)
    try:
        for gates, mode in ((brick, jit.ALWAYS), (rand, jit.OFF)):
            jit.set_jit(mode)
            psi = bf.random_state(n, rng)
            want = oracle.execute(psi, n, gates)
            launches = jit.jit_stats()["launches"]
            assert np.max(np.abs(run(psi, n, gates, lq.C128, abi.EXEC_FUSED) - want)) <= 1e-12
            assert np.max(np.abs(run(psi, n, gates, lq.C64, abi.EXEC_FUSED) - want)) <= 2e-6
            ran = jit.jit_stats()["launches"] - launches
            assert (ran > 0) == (mode == jit.ALWAYS), (mode, ran)
    finally:
        modes.restore_defaults()
