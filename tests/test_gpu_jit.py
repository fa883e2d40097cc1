"""Tile passes specialised at run time (hipRTC) against the oracle and the
generic tile kernel: complex128 within 1e-12 of the oracle and 1e-13 of the
generic kernel (the same arithmetic), complex64 within 2e-6 of the oracle;
through the fused executor, the batched parameter shift (eager, captured and
replayed from hipGraphs) and the batched adjoint gradient; and the
specialised kernels are the ones launched."""
import math

import numpy as np
import pytest

import oracle
import logosq_amd as lq
from logosq_amd import abi, algorithms
from logosq_amd import jit
from tests import bruteforce as bf
from tests import modes
from tests.test_gpu_fused import run
from tests.test_planner import random_circuit

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture
def jit_mode():
    yield jit.set_jit
    modes.restore_defaults()


def _launched(before):
    after = jit.jit_stats()
    assert after["failed"] == before["failed"], "hipRTC failed on a pass kernel"
    return after["launches"] - before["launches"]


@pytest.mark.parametrize("n", [8, 12, 16, 18])
def test_jit_random_circuits(hip, jit_mode, n):
    say("jit random n=%d" % n)
    rng = np.random.default_rng(900 + n)
    gates = random_circuit(n, 150, rng, with_dense=n <= 12)
    psi = bf.random_state(n, rng)
    want = oracle.execute(psi, n, gates)
    jit_mode(jit.OFF)
    gen = run(psi, n, gates, lq.C128, abi.EXEC_FUSED)
    gen64 = run(psi, n, gates, lq.C64, abi.EXEC_FUSED)
    jit_mode(jit.ALWAYS)
    s0 = jit.jit_stats()
    got = run(psi, n, gates, lq.C128, abi.EXEC_FUSED)
    got64 = run(psi, n, gates, lq.C64, abi.EXEC_FUSED)
    assert _launched(s0) > 0
    assert np.max(np.abs(got - want)) <= 1e-12
    assert np.max(np.abs(got - gen)) <= 1e-13
    assert np.max(np.abs(got64 - want)) <= 2e-6
    assert np.max(np.abs(got64 - gen64)) <= 1e-6


def sign_heavy_circuit(n, layers, rng):
    """RY / RZ layers with CZ rings, Z gates and CPs between them: Z and CZ
    become sign flips in the specialised kernels (controls inside the
    subcube: compile-time negations; outside: per-thread flags), beside
    phase-table runs and thread phases of the other diagonals."""
    from tests.test_planner import G
    out = []
    for _ in range(layers):
        for q in range(n):
            out.append(G(abi.RY, [q], [rng.uniform(0, 6.28)]))
        for q in range(n):
            out.append(G(abi.CZ, [q, (q + 1) % n]))
        for q in rng.choice(n, n // 3, replace=False):
            out.append(G(abi.Z, [int(q)]))
        for q in range(n):
            out.append(G(abi.RZ, [q], [rng.uniform(0, 6.28)]))
        a, b = rng.choice(n, 2, replace=False)
        out.append(G(abi.CP, [int(a), int(b)], [rng.uniform(0, 6.28)]))
        a, b = rng.choice(n, 2, replace=False)
        out.append(G(abi.CZ, [int(a), int(b)]))
    return out


@pytest.mark.parametrize("n", [13, 16])
def test_jit_sign_ops(hip, jit_mode, n):
    say("jit sign ops n=%d" % n)
    rng = np.random.default_rng(4200 + n)
    gates = sign_heavy_circuit(n, 6, rng)
    psi = bf.random_state(n, rng)
    want = oracle.execute(psi, n, gates)
    jit_mode(jit.OFF)
    gen = run(psi, n, gates, lq.C128, abi.EXEC_FUSED)
    jit_mode(jit.ALWAYS)
    s0 = jit.jit_stats()
    got = run(psi, n, gates, lq.C128, abi.EXEC_FUSED)
    got64 = run(psi, n, gates, lq.C64, abi.EXEC_FUSED)
    assert _launched(s0) > 0
    assert np.max(np.abs(got - want)) <= 1e-12
    assert np.max(np.abs(got - gen)) <= 1e-13
    assert np.max(np.abs(got64 - want)) <= 2e-6


def test_jit_brickwork_22(hip, jit_mode):
    """Brickwork at 22 qubits (1024 tiles per pass): the repeat policy runs
    the generic kernel the first time and the specialised kernels from the
    second run on."""
    say("jit brickwork 22")
    n = 22
    d = algorithms.brickwork_circuit(n, 6, 4).descriptors()
    psi = oracle.zero_state(n)
    want = oracle.execute(psi, n, d)
    jit_mode(jit.REPEAT)
    s0 = jit.jit_stats()
    first = run(psi, n, d, lq.C128, abi.EXEC_FUSED)
    s1 = jit.jit_stats()
    assert s1["launches"] == s0["launches"]
    second = run(psi, n, d, lq.C128, abi.EXEC_FUSED)
    assert _launched(s1) > 0
    assert np.max(np.abs(first - want)) <= 1e-12
    assert np.max(np.abs(second - want)) <= 1e-12
    assert np.max(np.abs(second - first)) <= 1e-13


def test_jit_batched_parameter_shift_with_graphs(hip, jit_mode):
    """Batched parameter shift at 14 qubits: generic eager run, then the
    specialised kernels eager, captured into a hipGraph and replayed."""
    say("jit parameter shift 14")
    n = 14
    ans = lq.ry_rz_cz_ring_ansatz(n, 2)
    obs = lq.heisenberg_ring(n, 0.4, -0.8, 0.3)
    x = np.random.default_rng(77).uniform(0, 2 * math.pi, ans.num_parameters())
    ps = lq.ParameterShift(batched=True)
    jit_mode(jit.OFF)
    e_ref, g_ref = ps.compute_gradient(ans, obs, x, energy=True)
    jit_mode(jit.ALWAYS)
    s0 = jit.jit_stats()
    for rep in range(3):
        e, g = ps.compute_gradient(ans, obs, x, energy=True)
        assert abs(e - e_ref) <= 1e-10, rep
        assert np.max(np.abs(np.array(g) - np.array(g_ref))) <= 1e-10, rep
    assert _launched(s0) > 0


def test_jit_adjoint_batch(hip, jit_mode):
    say("jit adjoint batch 16")
    n = 16
    ans = lq.ry_rz_cz_ring_ansatz(n, 2)
    rng = np.random.default_rng(78)
    J = rng.uniform(-1, 1, (5, 3))
    obs = [lq.heisenberg_ring(n, *map(float, j)) for j in J]
    x = rng.uniform(0, 2 * math.pi, (5, ans.num_parameters()))
    adj = lq.AdjointGradient()
    jit_mode(jit.OFF)
    e_ref, g_ref = adj.compute_gradient_batch(ans, obs, x, energy=True)
    jit_mode(jit.ALWAYS)
    s0 = jit.jit_stats()
    e, g = adj.compute_gradient_batch(ans, obs, x, energy=True)
    assert _launched(s0) > 0
    assert np.max(np.abs(e - e_ref)) <= 1e-10
    assert np.max(np.abs(g - g_ref)) <= 1e-10
