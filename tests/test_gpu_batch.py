"""GPU parity of the batched paths against the oracle: all 2P+1
parameter-shift circuits in one batch (energy and gradient <= 1e-10 for
complex128, <= 1e-5 for complex64 with f64 reductions), and batches of
circuits with uneven qubit counts and their own Pauli sums."""
import math
import os

import numpy as np
import pytest

import oracle
import logosq_amd as lq
from logosq_amd import abi, algorithms

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("n,layers", [(4, 2), (9, 2), (13, 2), (15, 1)])
def test_batched_param_shift_matches_oracle(hip, n, layers):
    say("batched ps n=%d L=%d" % (n, layers))
    rng = np.random.default_rng(900 + n)
    ans = lq.ry_rz_cz_ring_ansatz(n, layers)
    obs = lq.heisenberg_ring(n, *rng.uniform(-1, 1, 3))
    x = rng.uniform(0, 2 * math.pi, ans.num_parameters())
    e, g = lq.ParameterShift(batched=True).compute_gradient(ans, obs, x, energy=True)
    codes, coeffs = obs.codes()
    t = ans.template()
    assert abs(e - oracle.energy(n, t, x, codes, coeffs)) <= 1e-10
    gref = oracle.param_shift(n, t, x, codes, coeffs)
    assert np.max(np.abs(np.array(g) - gref)) <= 1e-10
    e64, g64 = lq.ParameterShift(batched=True, dtype=lq.C64).compute_gradient(ans, obs, x, energy=True)
    assert abs(e64 - e) <= 1e-5 and np.max(np.abs(np.array(g64) - gref)) <= 1e-5


def test_batched_equals_sequential_with_y_terms(hip):
    """Random Pauli sum with odd-Y terms (reference convention) on an
    EfficientSU2-style template: batched == sequential == oracle."""
    say("batched vs sequential")
    n = 7
    rng = np.random.default_rng(17)
    ans = lq.EfficientSU2Ansatz(n, 2)
    obs = lq.PauliObservable(n)
    for _ in range(12):
        obs.add_term(lq.PauliTerm(rng.normal(), rng.integers(0, 4, n)))
    x = rng.uniform(0, 2 * math.pi, ans.num_parameters())
    g_seq = lq.ParameterShift().compute_gradient(ans, obs, x)
    g_bat = lq.ParameterShift(batched=True).compute_gradient(ans, obs, x)
    codes, coeffs = obs.codes()
    gref = oracle.param_shift(n, ans.template(), x, codes, coeffs)
    assert np.max(np.abs(np.array(g_seq) - gref)) <= 1e-10
    assert np.max(np.abs(np.array(g_bat) - gref)) <= 1e-10


@pytest.mark.parametrize("dtype,tol", [(lq.C128, 1e-10), (lq.C64, 1e-5)])
def test_uneven_batch_matches_oracle(hip, dtype, tol):
    say("uneven batch dtype=%d" % dtype)
    items = algorithms.uneven_batch_workload(count=24, nmin=3, nmax=14, layers=2, terms=8, seed=11)
    got = lq.batch_expectations(items, dtype)
    for (n, gates, codes, coeffs), e in zip(items, got):
        psi = oracle.execute(oracle.zero_state(n), n, gates)
        want = oracle.pauli_sum(psi, n, codes, coeffs)
        assert abs(e - want) <= tol, (n, e, want)


@pytest.mark.parametrize("dtype,tol", [(lq.C128, 1e-10), (lq.C64, 1e-5)])
def test_uneven_batch_one_sequence_every_call(hip, dtype, tol):
    """One uneven batch through every stage of its launch sequence, each call
    against the oracle: circuits with no gates (|0...0> written by a fill),
    fewer than 8 qubits (generic pass kernel, fill), a renormalising
    MatrixGate (per-instance normalize step, not captured), and structures
    that recur, whose first pass is specialised to build |0...0> in LDS once
    compiled (second call) and whose sequence is captured (third) and
    replayed (fourth)."""
    from tests.test_oracle import G
    say("uneven batch, every call, dtype=%d" % dtype)
    rng = np.random.default_rng(23)
    items = list(algorithms.uneven_batch_workload(count=20, nmin=6, nmax=15, layers=2, terms=6, seed=29))
    for n in (9, 12):
        items.append((n, [], rng.integers(0, 4, (5, n)).astype(np.uint8), rng.normal(size=5)))
    m = np.linalg.qr(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))[0] * 1.2
    dense = [G(abi.RY, [q], [0.3 + q]) for q in range(10)] + [G(abi.MATRIX1, [4], (), m.reshape(-1),
                                                                  abi.GATE_FLAG_RENORMALIZE)]
    dense += [G(abi.CNOT, [q, q + 1]) for q in range(9)]
    with_dense = items + [(10, dense, rng.integers(0, 4, (7, 10)).astype(np.uint8), rng.normal(size=7))]
    want = []
    for n, gates, codes, coeffs in with_dense:
        psi = oracle.execute(oracle.zero_state(n), n, gates)
        want.append(oracle.pauli_sum(psi, n, codes, coeffs))
    want = np.array(want)
    results = []
    for call in range(4):
        got = lq.batch_expectations(items, dtype)
        err = np.abs(got - want[:len(items)])
        say("  call %d: max |dE| %.3e" % (call, err.max()))
        assert err.max() <= tol, (call, int(np.argmax(err)))
        results.append(got)
    # eager, captured and replayed runs of the same kernels agree bit for bit
    assert np.array_equal(results[1], results[2]) and np.array_equal(results[2], results[3])
    for call in range(2):
        got = lq.batch_expectations(with_dense, dtype)
        err = np.abs(got - want)
        say("  with a MatrixGate, call %d: max |dE| %.3e" % (call, err.max()))
        assert err.max() <= tol, (call, int(np.argmax(err)))


def test_uneven_batch_split_into_jobs(hip):
    """A memory budget smaller than the batch (LQ_BATCH_BUDGET_MIB=1): the
    circuits run as several launch sequences, a structure's instances split
    across pieces, states over the budget alone; every energy still equals
    the oracle's, and the same as one unsplit call."""
    say("uneven batch split by a 1 MiB budget")
    items = algorithms.uneven_batch_workload(count=30, nmin=8, nmax=17, layers=2, terms=6, seed=31)
    whole = lq.batch_expectations(items, lq.C128)
    os.environ["LQ_BATCH_BUDGET_MIB"] = "1"
    try:
        split = [lq.batch_expectations(items, lq.C128) for _ in range(3)]
    finally:
        del os.environ["LQ_BATCH_BUDGET_MIB"]
    want = np.array([oracle.pauli_sum(oracle.execute(oracle.zero_state(n), n, gates), n, codes, coeffs)
                     for n, gates, codes, coeffs in items])
    for e in [whole] + split:
        assert np.max(np.abs(e - want)) <= 1e-10
    assert np.max(np.abs(split[0] - whole)) <= 1e-13


@pytest.mark.parametrize("n,layers", [(3, 2), (8, 3), (14, 2)])
def test_adjoint_gradient_matches_param_shift(hip, n, layers):
    """The adjoint gradient equals the oracle's parameter-shift gradient
    (<= 1e-10), on the ring ansatz and on EfficientSU2 with a random Pauli sum
    holding odd-Y terms (reference convention)."""
    say("adjoint n=%d" % n)
    rng = np.random.default_rng(700 + n)
    for ans, obs in ((lq.ry_rz_cz_ring_ansatz(n, layers), lq.heisenberg_ring(n, *rng.uniform(-1, 1, 3))),
                     (lq.EfficientSU2Ansatz(n, layers), None)):
        if obs is None:
            obs = lq.PauliObservable(n)
            for _ in range(10):
                obs.add_term(lq.PauliTerm(rng.normal(), rng.integers(0, 4, n)))
        x = rng.uniform(0, 2 * math.pi, ans.num_parameters())
        e, g = lq.AdjointGradient().compute_gradient(ans, obs, x, energy=True)
        codes, coeffs = obs.codes()
        t = ans.template()
        assert abs(e - oracle.energy(n, t, x, codes, coeffs)) <= 1e-10
        assert np.max(np.abs(np.array(g) - oracle.param_shift(n, t, x, codes, coeffs))) <= 1e-10


def test_vqe_adjoint_equals_parameter_shift_trajectory(hip):
    """VQE::run with the adjoint gradient follows the parameter-shift
    trajectory (oracle) to rounding."""
    say("vqe adjoint")
    n = 5
    ans = lq.ry_rz_cz_ring_ansatz(n, 2)
    obs = lq.heisenberg_ring(n, 0.3, 0.9, -0.4)
    x0 = np.linspace(0.1, 1.9, ans.num_parameters())
    res = lq.VQE(ans, obs, lq.AdjointGradient(), lq.Adam(0.05, 12)).run(x0)
    codes, coeffs = obs.codes()
    xo, eo, ho, conv, final = oracle.vqe_run(n, ans.template(), ans.num_parameters(), codes, coeffs, x0, 1, 0.05, 12)
    assert np.max(np.abs(np.array(res.convergence_history) - ho)) <= 1e-9
    assert abs(res.ground_state_energy - eo) <= 1e-9


def _mixed_template(n, layers):
    """RX/RY/RZ parameters between fixed gates of most kinds (H, S, T, the
    reference's Y, CNOT, SWAP, CPHASE, U3, RXX in the quirk orientation,
    Toffoli, CRY) and a CZ ring."""
    t, p = [], 0
    for layer in range(layers):
        for q in range(n):
            t.append(abi.make_gate([abi.RX, abi.RY, abi.RZ][(q + layer) % 3], [q], [0.0], [p]))
            p += 1
        t += [abi.make_gate(abi.H, [0]), abi.make_gate(abi.S, [1]), abi.make_gate(abi.T, [n - 1]),
              abi.make_gate(abi.Y, [n // 2]), abi.make_gate(abi.CNOT, [0, n - 1]),
              abi.make_gate(abi.SWAP, [1, n - 2]), abi.make_gate(abi.CPHASE, [n - 1, 0], [0.7]),
              abi.make_gate(abi.U3, [2], [0.3, 0.5, 0.9]), abi.make_gate(abi.RXX, [2, 1], [0.4]),
              abi.make_gate(abi.TOFFOLI, [0, 1, n - 1]), abi.make_gate(abi.CRY, [n - 1, 1], [0.8])]
        t += [abi.make_gate(abi.CZ, [q, (q + 1) % n]) for q in range(n)]
    return t, p


def _adjoint(n, dtype, t, x, codes, coeffs):
    import ctypes
    arr = abi.gate_array(t)
    x = np.ascontiguousarray(x, dtype=np.float64)
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
    g = np.zeros(len(x))
    e = ctypes.c_double()
    pd = ctypes.POINTER(ctypes.c_double)
    lq._lib.call("lq_adjoint_gradient", n, dtype, arr, len(t), x.ctypes.data_as(pd), len(x),
                 codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), coeffs.ctypes.data_as(pd), len(coeffs),
                 abi.PAULI_REFERENCE, ctypes.byref(e), g.ctypes.data_as(pd))
    return e.value, g


@pytest.mark.parametrize("n,dtype", [(4, lq.C128), (9, lq.C128), (13, lq.C128), (13, lq.C64)])
def test_adjoint_mixed_gates_matches_param_shift(hip, n, dtype):
    """The fused backward sweep (inverse ops of every kind, tile passes with
    high bits at n = 13) against the oracle's parameter shift: 1e-10 for
    complex128, 1e-5 for complex64."""
    say("adjoint mixed n=%d dtype=%d" % (n, dtype))
    rng = np.random.default_rng(1300 + n)
    t, P = _mixed_template(n, 2)
    x = rng.uniform(0, 2 * math.pi, P)
    codes = rng.integers(0, 4, size=(12, n)).astype(np.uint8)
    coeffs = rng.normal(size=12)
    e, g = _adjoint(n, dtype, t, x, codes, coeffs)
    tol = 1e-10 if dtype == lq.C128 else 1e-5
    assert abs(e - oracle.energy(n, t, x, codes, coeffs)) <= tol
    want = oracle.param_shift(n, t, x, codes, coeffs)
    assert np.max(np.abs(g - want)) <= tol
    assert np.max(np.abs(want)) > 1e-3  # the gradient is not trivially zero


@pytest.mark.parametrize("n,layers", [(5, 2), (12, 2)])
def test_adjoint_batch_equals_single_and_oracle(hip, n, layers):
    """Batched adjoint gradients (many couplings and parameter sets of one
    ansatz) equal the single-instance ones and the oracle's parameter shift."""
    say("adjoint batch n=%d" % n)
    rng = np.random.default_rng(4100 + n)
    ans = lq.ry_rz_cz_ring_ansatz(n, layers)
    P = ans.num_parameters()
    Js = rng.uniform(-1, 1, (5, 3))
    obs = [lq.heisenberg_ring(n, *J) for J in Js]
    X = rng.uniform(0, 2 * math.pi, (5, P))
    adj = lq.AdjointGradient()
    E, G = adj.compute_gradient_batch(ans, obs, X, energy=True)
    for i in range(5):
        e1, g1 = adj.compute_gradient(ans, obs[i], X[i], energy=True)
        assert abs(E[i] - e1) <= 1e-14 and np.max(np.abs(G[i] - np.array(g1))) <= 1e-14
    t = ans.template()
    for i in (0, 4):
        codes, coeffs = obs[i].codes()
        assert abs(E[i] - oracle.energy(n, t, X[i], codes, coeffs)) <= 1e-10
        assert np.max(np.abs(G[i] - oracle.param_shift(n, t, X[i], codes, coeffs))) <= 1e-10
    E64, G64 = lq.AdjointGradient(lq.C64).compute_gradient_batch(ans, obs, X, energy=True)
    assert np.max(np.abs(E64 - E)) <= 1e-5 and np.max(np.abs(G64 - G)) <= 1e-5


def test_vqe_adam_batch_equals_single_runs(hip):
    """Lock-step VQE over several couplings: each instance's history,
    optimum and stopping match its own VQE::run (adjoint gradient), and one
    matches the oracle's parameter-shift VQE."""
    say("vqe adam batch")
    n = 6
    ans = lq.ry_rz_cz_ring_ansatz(n, 2)
    rng = np.random.default_rng(4200)
    Js = rng.uniform(-1, 1, (4, 3))
    obs = [lq.heisenberg_ring(n, *J) for J in Js]
    X0 = rng.uniform(0, 2 * math.pi, (4, ans.num_parameters()))
    for lr, it, tol in ((0.05, 40, 1e-6), (0.1, 60, 2e-3)):
        res, conv = lq.vqe_adam_batch(ans, obs, X0, lq.Adam(lr, it).with_tolerance(tol))
        iters = []
        for i in range(4):
            single = lq.VQE(ans, obs[i], lq.AdjointGradient(), lq.Adam(lr, it).with_tolerance(tol)).run(X0[i])
            assert len(single.convergence_history) == len(res[i].convergence_history)
            assert np.max(np.abs(np.array(single.convergence_history) - res[i].convergence_history)) <= 1e-12
            assert np.max(np.abs(np.array(single.optimal_parameters) - res[i].optimal_parameters)) <= 1e-12
            assert abs(single.ground_state_energy - res[i].ground_state_energy) <= 1e-12
            iters.append(len(res[i].convergence_history))
        if tol > 1e-3:
            assert min(iters) < it and conv.any()  # instances stop on their own
    codes, coeffs = obs[0].codes()
    res, _ = lq.vqe_adam_batch(ans, obs, X0, lq.Adam(0.05, 25))
    xo, eo, ho, co, _ = oracle.vqe_run(n, ans.template(), ans.num_parameters(), codes, coeffs, X0[0], 1, 0.05, 25)
    assert np.max(np.abs(np.array(res[0].convergence_history) - ho)) <= 1e-9
    assert abs(res[0].ground_state_energy - eo) <= 1e-9


def test_batched_param_shift_shared_prefix_edges(hip):
    """Shared prefixes at their edges: parameter 0 drives the first and the
    last gate, parameter 1 no gate (its shifted circuits equal the base one
    and join only after the last step), parameter 2 only the last gate, the
    rest a brickwork in between.  Batched against sequential and the
    oracle."""
    print("batched parameter shift, shared-prefix edges", flush=True)
    n = 16  # wider than one tile: the plan has several steps

    def build(p):
        c = lq.Circuit(n)
        c.ry(0, p[0])
        for layer in range(3):
            for q in range(n):
                c.rx(q, p[3 + layer * n + q])
            for q in range(layer % 2, n - 1, 2):
                c.cnot(q, q + 1)
        c.rz(5, p[2])
        c.ry(n - 1, p[0])
        return c

    ans = lq.ParameterizedCircuit(n, 3 + 3 * n, build)
    assert algorithms_passes(n, ans.template()) >= 2
    obs = lq.heisenberg_ring(n, 0.5, -0.6, 0.7)
    x = np.random.default_rng(92).uniform(0, 2 * math.pi, ans.num_parameters())
    e_b, g_b = lq.ParameterShift(batched=True).compute_gradient(ans, obs, x, energy=True)
    e_s, g_s = lq.ParameterShift(batched=False).compute_gradient(ans, obs, x, energy=True)
    assert abs(e_b - e_s) <= 1e-12
    assert np.max(np.abs(np.array(g_b) - np.array(g_s))) <= 1e-12
    assert abs(g_b[1]) <= 1e-12  # the unused parameter
    codes, coeffs = obs.codes()
    want = oracle.param_shift(n, ans.template(), x, codes, coeffs)
    assert np.max(np.abs(np.array(g_b) - want)) <= 1e-10


def algorithms_passes(n, gates):
    import ctypes
    lib = lq.load()
    order = (ctypes.c_int * len(gates))()
    npass = ctypes.c_int()
    assert lib.lq_plan_fused(n, abi.LQ_C128, abi.gate_array(gates), len(gates), 0, 0, order, None,
                             ctypes.byref(npass)) == 0
    return npass.value



def test_vqe_run_to_target_follows_vqe_run(hip):
    """lq_vqe_run_to_target: the same iterations as VQE::run (Adam,
    src/optimization/optimizer.rs:139-223) up to the first energy at or below
    the target, where it stops; with an unreachable target it is VQE::run."""
    say("vqe run to target")
    n = 6
    ans = lq.ry_rz_cz_ring_ansatz(n, 2)
    obs = lq.heisenberg_ring(n, 0.4, -0.8, 0.3)
    x0 = np.linspace(0.2, 2.2, ans.num_parameters())
    for grad in (lq.AdjointGradient(), lq.ParameterShift(batched=True), lq.ParameterShift()):
        full = lq.VQE(ans, obs, grad, lq.Adam(0.05, 40)).run(x0)
        h = np.array(full.convergence_history)
        assert len(h) == 40
        target = float(h[9])
        first = int(np.flatnonzero(h <= target)[0])
        r, reached = lq.VQE(ans, obs, grad, lq.Adam(0.05, 40)).run_to_target(x0, target)
        assert reached and len(r.convergence_history) == first + 1
        assert np.array_equal(np.array(r.convergence_history), h[:first + 1])
        r2, reached2 = lq.VQE(ans, obs, grad, lq.Adam(0.05, 40)).run_to_target(x0, -1e9)
        assert not reached2 and np.array_equal(np.array(r2.convergence_history), h)
        assert np.max(np.abs(np.array(r2.optimal_parameters) - np.array(full.optimal_parameters))) == 0.0
