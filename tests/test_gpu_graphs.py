"""Batched parameter-shift gradients and uneven batches replayed from
hipGraphs give the same results as the eager launch sequence, replay after
replay, with other eager work and synchronisations in between."""
import ctypes
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import logosq_amd as lq
from logosq_amd import algorithms

pytestmark = pytest.mark.gpu


def _stats():
    c, r = ctypes.c_longlong(), ctypes.c_longlong()
    lq._lib.call("lq_batch_graph_stats", ctypes.byref(c), ctypes.byref(r))
    return c.value, r.value


def _other_work():
    st = lq.State(12)
    c = algorithms.brickwork_circuit(12, 3, 7)
    c.execute(st)
    lq._lib.call("lq_synchronize")
    return st.norm_sqr()


def test_param_shift_graph_replays_match_eager(hip):
    print("graphs: parameter shift", flush=True)
    n = 10
    ans = lq.ry_rz_cz_ring_ansatz(n, 2)
    obs = lq.heisenberg_ring(n, 0.4, -0.8, 0.3)
    rng = np.random.default_rng(61)
    xs = [rng.uniform(0, 2 * math.pi, ans.num_parameters()) for _ in range(3)]
    ps = lq.ParameterShift(batched=True)
    try:
        lq._lib.call("lq_set_batch_graphs", 0)
        eager = [ps.compute_gradient(ans, obs, x, energy=True) for x in xs]
        lq._lib.call("lq_set_batch_graphs", 1)
        c0, r0 = _stats()
        for rep in range(4):
            for x, (e_ref, g_ref) in zip(xs, eager):
                e, g = ps.compute_gradient(ans, obs, x, energy=True)
                assert e == e_ref and np.array_equal(np.array(g), np.array(g_ref)), rep
                _other_work()
        c1, r1 = _stats()
        assert c1 - c0 >= 1 and r1 - r0 >= 8
    finally:
        lq._lib.call("lq_set_batch_graphs", 1)


def test_uneven_batch_graph_replays_match_eager(hip):
    print("graphs: uneven batch", flush=True)
    items = algorithms.uneven_batch_workload(24, 6, 11, 2, 8, 9)
    try:
        lq._lib.call("lq_set_batch_graphs", 0)
        # the second sighting compiles the specialised passes; the eager
        # reference is taken with the kernels the replays will use
        for _ in range(2):
            lq.batch_expectations(items, lq.C128)
        ref = lq.batch_expectations(items, lq.C128)
        lq._lib.call("lq_set_batch_graphs", 1)
        c0, r0 = _stats()
        for rep in range(4):
            e = lq.batch_expectations(items, lq.C128)
            assert np.array_equal(e, ref), rep
            _other_work()
        c1, r1 = _stats()
        assert c1 > c0 and r1 > r0
    finally:
        lq._lib.call("lq_set_batch_graphs", 1)


_FRESH = r"""
import ctypes, math, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import logosq_amd as lq
lq.load()
ans = lq.ry_rz_cz_ring_ansatz(10, 2)
obs = lq.heisenberg_ring(10, 0.4, -0.8, 0.3)
x = np.random.default_rng(5).uniform(0, 2 * math.pi, ans.num_parameters())
ps = lq.ParameterShift(batched=True)
gs = [np.array(ps.compute_gradient(ans, obs, x)) for _ in range(5)]
c, r = ctypes.c_longlong(), ctypes.c_longlong()
lq._lib.call("lq_batch_graph_stats", ctypes.byref(c), ctypes.byref(r))
print("captures", c.value, "replays", r.value, "equal", all(np.array_equal(g, gs[1]) for g in gs[1:]))
"""


def test_batched_gradient_captured_in_a_fresh_process(hip):
    """The batched parameter shift as the first call of a process: the
    library stream is created by that call (not left to the legacy null
    stream, where a capture is refused), so the third call captures the
    launch sequence and the later ones replay it, with identical results."""
    print("graphs: fresh process", flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", _FRESH, root], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr[-2000:]
    line = out.stdout.strip().splitlines()[-1]
    print("  " + line, flush=True)
    f = line.split()
    assert int(f[1]) >= 1 and int(f[3]) >= 2 and f[5] == "True", line
