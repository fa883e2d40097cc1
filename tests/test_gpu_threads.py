"""Calls from several host threads on distinct handles: every entry point
takes one library lock, so concurrent circuit executions and batched
parameter-shift gradients (which share the library stream, scratch buffers
and plan caches) give the same results as the same calls made one after
another, and the circuits match the oracle (complex128 <= 1e-12)."""
import math
import threading

import numpy as np
import pytest

import oracle
import logosq_amd as lq
from logosq_amd import abi
from tests import bruteforce as bf
from tests.test_gpu_fused import run
from tests.test_planner import random_circuit

pytestmark = pytest.mark.gpu

NTHREADS = 4


def _in_threads(work):
    out = [None] * NTHREADS
    errs = []

    def body(k):
        try:
            out[k] = work(k)
        except BaseException as e:  # reported by the main thread
            errs.append((k, e))

    ts = [threading.Thread(target=body, args=(k,)) for k in range(NTHREADS)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in ts), "a worker thread did not finish"
    assert not errs, errs
    return out


def test_threads_execute_distinct_states(hip):
    print("threads: fused execute n=14", flush=True)
    n = 14
    cases = []
    for k in range(NTHREADS):
        rng = np.random.default_rng(1300 + k)
        gates = random_circuit(n, 120, rng, with_dense=False)
        psi = bf.random_state(n, rng)
        cases.append((psi, gates, oracle.execute(psi, n, gates)))

    def work(k):
        psi, gates, _ = cases[k]
        return [run(psi, n, gates, lq.C128, abi.EXEC_FUSED) for _ in range(3)]

    got = _in_threads(work)
    for k in range(NTHREADS):
        for v in got[k]:
            assert np.max(np.abs(v - cases[k][2])) <= 1e-12, k


def test_threads_batched_parameter_shift(hip):
    print("threads: batched parameter shift n=12", flush=True)
    n = 12
    ans = lq.ry_rz_cz_ring_ansatz(n, 2)
    rng = np.random.default_rng(1310)
    J = rng.uniform(-1, 1, (NTHREADS, 3))
    obs = [lq.heisenberg_ring(n, *map(float, j)) for j in J]
    xs = rng.uniform(0, 2 * math.pi, (NTHREADS, ans.num_parameters()))
    ps = lq.ParameterShift(batched=True)
    ref = [ps.compute_gradient(ans, obs[k], xs[k], energy=True) for k in range(NTHREADS)]

    def work(k):
        return [ps.compute_gradient(ans, obs[k], xs[k], energy=True) for _ in range(3)]

    got = _in_threads(work)
    for k in range(NTHREADS):
        e_ref, g_ref = ref[k]
        for e, g in got[k]:
            assert abs(e - e_ref) <= 1e-10, k
            assert np.max(np.abs(np.array(g) - np.array(g_ref))) <= 1e-10, k
