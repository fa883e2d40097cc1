"""GPU parity: every kernel of the HIP path against the CPU oracle on the
same seeded inputs.  Tolerances: complex128 amplitudes <= 1e-12,
expectations and gradients <= 1e-10; complex64 amplitudes <= 2e-6
(single-precision rounding over the test circuits), its f64-accumulated
expectations <= 1e-5."""
import math

import numpy as np
import pytest

import oracle
import logosq_amd as lq
from logosq_amd import abi
from tests import bruteforce as bf
from tests import modes
from tests.test_oracle import all_gates, G, random_codes, ring_template, hea_template

pytestmark = pytest.mark.gpu

AMP_TOL = 1e-12
C64_TOL = 2e-6


def say(*a):
    print(*a, flush=True)


def gpu_apply(psi, n, gates, dtype=lq.C128, mode=abi.EXEC_PER_GATE):
    st = lq.State(n, dtype)
    st.set_vector(psi)
    if gates:
        lq._lib.call("lq_execute", st.handle, abi.gate_array(gates), len(gates), mode)
    return st.vector()


@pytest.mark.parametrize("n", [1, 2, 3, 5, 6, 7, 8, 9])
def test_every_gate_every_target(hip, n):
    say("gates n=%d" % n)
    rng = np.random.default_rng(1000 + n)
    psi = bf.random_state(n, rng)
    gates = all_gates(n, rng)
    if n > 6:  # keep the 3-qubit combinations bounded
        gates = [g for g in gates if g.kind != abi.TOFFOLI] + [
            G(abi.TOFFOLI, [a, b, c]) for a, b, c in [(0, n - 1, 3), (n - 1, 0, n - 2), (2, 5, 0), (n - 1, 1, 6 % (n - 1))]]
    for g in gates:
        got = gpu_apply(psi, n, [g])
        want = oracle.apply_gate(psi, n, g)
        err = np.max(np.abs(got - want))
        assert err <= AMP_TOL, (g.kind, list(g.qubits), err)


@pytest.mark.parametrize("unroll,nt", [(1, 0), (2, 0), (2, 1), (4, 1), (8, 0), (8, 1)])
def test_gate_kernel_variants(hip, unroll, nt):
    """Every per-gate kernel variant the autotuner can pick gives the oracle's
    result (complex128 and complex64)."""
    say("variant unroll=%d nt=%d" % (unroll, nt))
    lq._lib.call("lq_set_gate_kernel_config", unroll, nt)
    try:
        from tests.test_planner import random_circuit
        for n in (5, 9, 14):
            rng = np.random.default_rng(77 + n)
            gates = random_circuit(n, 60, rng)
            psi = bf.random_state(n, rng)
            want = oracle.execute(psi, n, gates)
            assert np.max(np.abs(gpu_apply(psi, n, gates) - want)) <= AMP_TOL
            assert np.max(np.abs(gpu_apply(psi, n, gates, lq.C64) - want)) <= C64_TOL
    finally:
        lq._lib.call("lq_set_gate_kernel_config", 1, 1)


def order_gates(n, rng):
    """Every gate kind at low, middle and top placements (pairs both ways
    round), so each per-gate kernel class runs at least once."""
    from tests.test_oracle import ONE_Q, ONE_Q_ROT, TWO_Q, TWO_Q_ROT
    out = []
    for q in sorted({0, 1, 3, 6, n // 2, n - 2, n - 1}):
        out += [G(k, [q]) for k in ONE_Q]
        out += [G(k, [q], [rng.uniform(0, 2 * math.pi)]) for k in ONE_Q_ROT]
        out.append(G(abi.U3, [q], rng.uniform(0, 2 * math.pi, 3)))
    for a, b in [(0, 1), (1, 0), (0, n - 1), (n - 1, 0), (2, 9), (9, 2), (n // 2, n // 2 + 1), (n - 2, n - 1),
                 (n - 1, n - 3), (4, 12)]:
        out += [G(k, [a, b]) for k in TWO_Q]
        out += [G(k, [a, b], [rng.uniform(0, 2 * math.pi)]) for k in TWO_Q_ROT]
    out += [G(abi.TOFFOLI, t) for t in ([0, 1, 2], [n - 1, 0, n // 2], [3, n - 2, 7])]
    return out


@pytest.mark.parametrize("n,dtype", [(17, lq.C128), (18, lq.C64), (20, lq.C128)])
def test_gate_unit_orders(hip, n, dtype):
    """The transposed work-unit order (lq_set_gate_unit_order) visits every
    unit once: with it on for every kernel class the state after a sequence
    of every gate kind equals the plain order's bit for bit, and the oracle's
    to the usual tolerance.  States of 2^17 amplitudes and more, so that the
    launches have the 1024 units the order needs."""
    say("unit orders n=%d dtype=%d" % (n, dtype))
    rng = np.random.default_rng(4100 + n)
    gates = order_gates(n, rng)
    psi = bf.random_state(n, rng)
    try:
        got = {}
        for classes in (0, abi.UNITS_ALL):
            lq._lib.call("lq_set_gate_unit_order", classes)
            got[classes] = gpu_apply(psi, n, gates, dtype)
        assert np.array_equal(got[0], got[abi.UNITS_ALL])
        want = oracle.execute(psi, n, gates)
        tol = AMP_TOL if dtype == lq.C128 else 1e-5
        assert np.max(np.abs(got[abi.UNITS_ALL] - want)) <= tol
    finally:
        lq._lib.call("lq_set_gate_unit_order", -1)


def test_autotune_report(hip):
    """lq_autotune at 16 qubits: the JSON report lists every gate variant,
    both unit orders and the tile geometries with a choice for each, and
    the configuration it leaves gives the oracle's results."""
    import ctypes
    import json
    say("autotune 16 qubits")
    buf = ctypes.create_string_buffer(16384)
    try:
        lq._lib.call("lq_autotune", 16, lq.C128, buf, len(buf))
        rep = json.loads(buf.value.decode())
        assert len(rep["gate_variants"]) == 8 and rep["gate_choice"]["unroll"] in (1, 2, 4, 8)
        assert [o["classes"] for o in rep["unit_orders"]] == [0, abi.UNITS_ALL]
        assert rep["unit_order_choice"] in (0, abi.UNITS_ALL)
        assert rep["tile_variants"] and all(v["ms"] > 0 for v in rep["tile_variants"])
        rng = np.random.default_rng(4200)
        gates = order_gates(16, rng)
        psi = bf.random_state(16, rng)
        want = oracle.execute(psi, 16, gates)
        assert np.max(np.abs(gpu_apply(psi, 16, gates) - want)) <= AMP_TOL
        assert np.max(np.abs(gpu_apply(psi, 16, gates, mode=abi.EXEC_FUSED) - want)) <= AMP_TOL
    finally:
        lq._lib.call("lq_set_gate_kernel_config", 1, 1)
        lq._lib.call("lq_set_gate_unit_order", -1)
        lq._lib.call("lq_set_fused_config", 0, 0, 0)


@pytest.mark.parametrize("n", [7, 12])
def test_gates_complex64(hip, n):
    say("c64 gates n=%d" % n)
    rng = np.random.default_rng(2000 + n)
    psi = bf.random_state(n, rng)
    for g in all_gates(min(n, 4), rng)[::3] + [G(abi.RX, [q], [0.3 * q + 0.1]) for q in range(n)] + \
            [G(abi.CNOT, [q, (q + 5) % n]) for q in range(n)]:
        got = gpu_apply(psi, n, [g], lq.C64)
        want = oracle.apply_gate(psi, n, g)
        assert np.max(np.abs(got - want)) <= C64_TOL, (g.kind, list(g.qubits))


def test_dense_builders_and_renormalise(hip):
    say("dense builders")
    n = 6
    rng = np.random.default_rng(7)
    psi = bf.random_state(n, rng) * 1.3
    gates = []
    for q in range(n):
        m = bf.u3(*rng.uniform(0, 6, 3))
        gates.append(G(abi.MATRIX1, [q], (), m.reshape(-1), abi.GATE_FLAG_RENORMALIZE))
    m4 = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))[0]
    gates.append(G(abi.MATRIX2, [4, 1], (), m4.reshape(-1), abi.GATE_FLAG_RENORMALIZE))
    got = gpu_apply(psi, n, gates)
    want = oracle.execute(psi, n, gates)
    assert np.max(np.abs(got - want)) <= AMP_TOL


def test_matrix_gate_dense_matvec(hip):
    """MatrixGate::apply (src/gates/mod.rs:77-93) on a full random matrix."""
    say("MatrixGate")
    n = 7
    rng = np.random.default_rng(8)
    psi = bf.random_state(n, rng)
    M = rng.normal(size=(1 << n, 1 << n)) + 1j * rng.normal(size=(1 << n, 1 << n))
    st = lq.State(n)
    st.set_vector(psi)
    lq.MatrixGate(M).apply(st)
    want = oracle.matrix_gate(psi, n, M)
    assert np.max(np.abs(st.vector() - want)) <= AMP_TOL


def test_state_api(hip):
    """Constructors, normalize (zero vector unchanged), inner products,
    probabilities (src/states/mod.rs:55-391; tests/test_normalize.rs,
    tests/test_parallel.rs:75-141)."""
    say("state api")
    for dtype, tol in ((lq.C128, 1e-15), (lq.C64, 1e-7)):
        s = lq.State.one_state(3, dtype)
        assert abs(s.vector()[7] - 1) < tol
        p = lq.State.plus_state(4, dtype)
        assert np.max(np.abs(p.vector() - 0.25)) < tol
        b = lq.State.bell_state(dtype)
        assert abs(b.vector()[0] - 1 / math.sqrt(2)) < tol and abs(b.vector()[3] - 1 / math.sqrt(2)) < tol
    v = lq.State.new(np.array([3, 4j]))
    assert np.max(np.abs(v.vector() - np.array([0.6, 0.8j]))) < 1e-16
    z = lq.State(2)
    z.set_vector(np.zeros(4))
    assert z.normalize() == 0.0 and np.all(z.vector() == 0)
    with pytest.raises(lq.InvalidStateDimension):
        lq.State.new(np.ones(3))
    rng = np.random.default_rng(3)
    n = 11
    a, b2 = bf.random_state(n, rng), bf.random_state(n, rng)
    sa, sb = lq.State(n), lq.State(n)
    sa.set_vector(a * 2.5)
    sb.set_vector(b2)
    ip = sa.inner_product(sb)
    ref = np.zeros(2)
    want = np.vdot(a * 2.5, b2)
    assert abs(ip - want) < 1e-12
    norm = sa.normalize()
    assert abs(norm - 2.5) < 1e-12
    assert np.max(np.abs(sa.vector() - oracle.normalize(a * 2.5)[0])) <= 1e-15
    assert np.max(np.abs(sb.probabilities() - np.abs(b2) ** 2)) < 1e-15
    c = sb.clone()
    assert np.array_equal(c.vector(), sb.vector())


@pytest.mark.parametrize("n", [1, 2, 4, 7, 10, 14])
def test_pauli_expectations(hip, n):
    say("expectations n=%d" % n)
    rng = np.random.default_rng(3000 + n)
    psi = bf.random_state(n, rng)
    codes, coeffs = random_codes(n, 40, rng)
    codes[0, :] = 0  # identity term
    st = lq.State(n)
    st.set_vector(psi)
    obs = lq.PauliObservable(n)
    for r, c in zip(codes, coeffs):
        obs.add_term(lq.PauliTerm(c, r))
    got = obs.expectation(st)
    want = oracle.pauli_sum(psi, n, codes, coeffs)
    assert abs(got - want) <= 1e-10
    if n <= 7:
        phys = lq.PauliObservable(n, abi.PAULI_PHYSICAL)
        phys.terms = obs.terms
        assert abs(phys.expectation(st) - bf.expectation_physical(psi, n, codes, coeffs)) <= 1e-10
    s64 = lq.State(n, lq.C64)
    s64.set_vector(psi)
    assert abs(obs.expectation(s64) - want) <= 1e-5


@pytest.mark.parametrize("mode", [abi.EXEC_PER_GATE])
def test_qft_12_qubits(hip, mode):
    """Config: plain QFT on 12 qubits, complex128, N(0,1) input normalised;
    all 4096 amplitudes vs the oracle (<=1e-12) and vs sqrt(N) ifft."""
    say("qft12")
    n = 12
    psi0 = lq.algorithms.normal_state_vector(n, 1)
    st = lq.State.new(psi0)
    lq.algorithms.qft_circuit(n).execute(st, mode)
    got = st.vector()
    ref = oracle.execute(oracle.normalize(psi0)[0], n, oracle.qft_gates(n))
    assert np.max(np.abs(got - ref)) <= AMP_TOL
    closed = np.fft.ifft(oracle.normalize(psi0)[0]) * math.sqrt(1 << n)
    assert np.max(np.abs(got - closed)) <= 1e-12


@pytest.mark.parametrize("n,layers", [(4, 2), (6, 3), (8, 2)])
def test_param_shift_gradient(hip, n, layers):
    say("ps grad n=%d L=%d" % (n, layers))
    rng = np.random.default_rng(4000 + n)
    ans = lq.ry_rz_cz_ring_ansatz(n, layers)
    obs = lq.heisenberg_ring(n, *rng.uniform(-1, 1, 3))
    x = rng.uniform(0, 2 * math.pi, ans.num_parameters())
    e, g = lq.ParameterShift().compute_gradient(ans, obs, x, energy=True)
    codes, coeffs = obs.codes()
    t = ans.template()
    assert abs(e - oracle.energy(n, t, x, codes, coeffs)) <= 1e-10
    assert np.max(np.abs(np.array(g) - oracle.param_shift(n, t, x, codes, coeffs))) <= 1e-10


def test_gradient_reference_tests(hip):
    """tests/test_optimization_gradient.rs:61-88, 212-263 on the GPU path."""
    say("gradient known answers")
    ans = lq.HardwareEfficientAnsatz(1, 0)
    g = lq.ParameterShift().compute_gradient(ans, lq.PauliObservable.single_z(1, 0), [math.pi / 4])
    assert abs(g[0] + math.sin(math.pi / 4)) < 1e-6
    ans = lq.RealAmplitudesAnsatz(2, 1)
    x = [0.5, 1.0, 1.5, 0.3, 0.8, 0.2]
    g1 = lq.ParameterShift().compute_gradient(ans, lq.PauliObservable.single_z(2, 0), x)
    o2 = lq.PauliObservable(2)
    o2.add_term(lq.PauliTerm(2.0, [lq.Pauli.Z, lq.Pauli.I]))
    g2 = lq.ParameterShift().compute_gradient(ans, o2, x)
    assert np.max(np.abs(np.array(g2) - 2 * np.array(g1))) < 1e-10
    fd = lq.FiniteDifference(1e-5).compute_gradient(ans, lq.PauliObservable.single_z(2, 0), x)
    assert np.max(np.abs(np.array(fd) - np.array(g1))) < 1e-4


def test_vqe_matches_oracle_trajectory(hip):
    """VQE::run with Adam (src/optimization/vqe.rs:55-99): same iterates as
    the oracle's restatement."""
    say("vqe")
    n = 4
    ans = lq.ry_rz_cz_ring_ansatz(n, 2)
    obs = lq.heisenberg_ring(n, 0.4, -0.8, 0.3)
    x0 = np.linspace(0.2, 1.5, ans.num_parameters())
    vqe = lq.VQE(ans, obs, lq.ParameterShift(), lq.Adam(0.1, 15))
    vqe.verbose = False
    res = vqe.run(x0)
    codes, coeffs = obs.codes()
    xo, eo, ho, conv, final = oracle.vqe_run(n, ans.template(), ans.num_parameters(), codes, coeffs, x0, 1, 0.1, 15)
    assert abs(res.ground_state_energy - eo) <= 1e-10
    assert np.max(np.abs(np.array(res.convergence_history) - ho)) <= 1e-10
    assert np.max(np.abs(np.array(res.optimal_parameters) - xo)) <= 1e-9
    assert np.max(np.abs(res.final_state.vector() - final)) <= 1e-10


def test_optimizers_native(hip):
    """tests/test_optimization.rs:15-149 against the native optimisers."""
    say("optimizers")
    r = lq.GradientDescent(0.1, 100).minimize(lambda p: p[0] ** 2, lambda p: [2 * p[0]], [5.0])
    assert r.converged and abs(r.optimal_parameters[0]) < 0.1
    r = lq.Adam(0.1, 200).minimize(lambda p: (p[0] - 3) ** 2, lambda p: [2 * (p[0] - 3)], [10.0])
    assert r.converged and abs(r.optimal_parameters[0] - 3) < 1e-3
    xo, vo, ho, co = oracle.adam_minimize(lambda p: (p[0] - 3) ** 2, lambda p: [2 * (p[0] - 3)], [10.0], 0.1, 200)
    assert r.optimal_value == vo and r.num_iterations == len(ho)


def test_errors_and_panics(hip):
    """Panic strings and error variants the reference tests match
    (tests/test_circuits_dense.rs:50-60, tests/test_optimization.rs:288-303)."""
    say("errors")
    c = lq.Circuit(2)
    with pytest.raises(lq.InvalidQubitIndex, match=r"Invalid qubit index: 2 \(must be < 2\)"):
        c.add_operation(lq.x_gate(), [2], "X on 2")
    with pytest.raises(AssertionError, match="Qubit index out of bounds"):
        lq.PauliObservable.single_z(2, 3)
    ans = lq.HardwareEfficientAnsatz(2, 1)
    with pytest.raises(AssertionError, match="State qubit count must match ansatz"):
        ans.apply(lq.State(3), [0.1] * 4)
    with pytest.raises(lq.CircuitQubitMismatch):
        lq.Circuit(3).execute(lq.State(2))
    st = lq.State(3)
    bad = G(abi.CNOT, [1, 1])
    with pytest.raises(lq.LogosQError):
        lq._lib.call("lq_apply_gate", st.handle, __import__("ctypes").byref(bad))
    with pytest.raises(lq.InvalidQubitIndex):
        lq._lib.call("lq_apply_gate", st.handle, __import__("ctypes").byref(G(abi.RX, [5], [0.1])))


def test_large_state_round_trip(hip):
    """28 qubits (4 GiB c128): a brickwork layer and its inverse return |0..0>
    exactly up to rounding; the norm stays 1."""
    say("28q round trip")
    n = 28
    st = lq.State(n)
    fwd = lq.algorithms.brickwork_circuit(n, 2, 99)
    fwd.execute(st)
    assert abs(st.norm_sqr() - 1.0) < 1e-12
    inv = []
    for g in reversed(fwd.descriptors()):
        g2 = abi.LqGate()
        __import__("ctypes").pointer(g2)[0] = g
        if g2.kind in (abi.RX, abi.RY, abi.RZ):
            g2.theta[0] = -g2.theta[0]
        inv.append(g2)
    lq._lib.call("lq_execute", st.handle, abi.gate_array(inv), len(inv), abi.EXEC_PER_GATE)
    p0 = st.inner_product(lq.State(n))
    assert abs(p0 - 1.0) < 1e-12


@pytest.mark.parametrize("n", [3, 6, 9, 16])
def test_diagonal_groups_over_32_terms(hip, n):
    """Diagonal (I/Z-only) terms share one group; more than 32 of them run in
    several passes, and the lane/chunk split of the sign must hold for states
    shorter and longer than a wavefront's 64 amplitudes."""
    say("diagonal group n=%d" % n)
    rng = np.random.default_rng(4000 + n)
    psi = bf.random_state(n, rng)
    codes = np.where(rng.random((45, n)) < 0.5, 3, 0).astype(np.uint8)
    codes[:5, :] = np.array([[1, 2] * n][0][:n], dtype=np.uint8)  # a few flip terms as well
    coeffs = rng.normal(size=45)
    st = lq.State(n)
    st.set_vector(psi)
    obs = lq.PauliObservable(n)
    for r, c in zip(codes, coeffs):
        obs.add_term(lq.PauliTerm(c, r))
    assert abs(obs.expectation(st) - oracle.pauli_sum(psi, n, codes, coeffs)) <= 1e-10


@pytest.mark.parametrize("n", [9, 13, 16, 20])
def test_flip_groups_from_tiles(hip, n):
    """Local Hamiltonians (ring bonds XX, YY, ZZ, XY, YZ and 3-local
    strings, so flip groups share tiles) plus long random strings whose flip
    bits fit no tile: the LDS-tile flip kernel and the pair kernel together
    against the oracle, complex128 and complex64."""
    say("flip groups from tiles n=%d" % n)
    rng = np.random.default_rng(5100 + n)
    psi = bf.random_state(n, rng)
    rows, coeffs = [], []
    for q in range(n):
        for a, b in ((1, 1), (2, 2), (3, 3), (1, 2), (2, 3)):
            r = np.zeros(n, np.uint8)
            r[q], r[(q + 1) % n] = a, b
            rows.append(r)
            coeffs.append(rng.normal())
        r = np.zeros(n, np.uint8)
        r[q], r[(q + 2) % n], r[(q + 5) % n] = 2, 1, 3
        rows.append(r)
        coeffs.append(rng.normal())
    for _ in range(4):
        rows.append(rng.integers(0, 4, n).astype(np.uint8))
        coeffs.append(rng.normal())
    codes, coeffs = np.array(rows), np.array(coeffs)
    obs = lq.PauliObservable(n)
    for r, c in zip(codes, coeffs):
        obs.add_term(lq.PauliTerm(c, r))
    want = oracle.pauli_sum(psi, n, codes, coeffs)
    st = lq.State(n)
    st.set_vector(psi)
    s64 = lq.State(n, lq.C64)
    s64.set_vector(psi)
    try:
        for mode in (abi.EXPECT_TILES_ON, abi.EXPECT_TILES_OFF):
            lq._lib.call("lq_set_expect_tiles", mode)
            assert abs(obs.expectation(st) - want) <= 1e-10, mode
            assert abs(obs.expectation(s64) - want) <= 1e-5, mode
    finally:
        modes.restore_defaults()


def test_batched_gradient_flip_groups_from_tiles(hip):
    """The batched parameter shift with its flip groups from LDS tiles (every
    instance shares the sum) equals the per-group kernels."""
    say("batched gradient, tiled flip groups")
    n = 12
    ans = lq.ry_rz_cz_ring_ansatz(n, 2)
    obs = lq.heisenberg_ring(n, 0.6, -0.3, 0.8)
    x = np.random.default_rng(91).uniform(0, 2 * math.pi, ans.num_parameters())
    ps = lq.ParameterShift(batched=True)
    try:
        lq._lib.call("lq_set_expect_tiles", abi.EXPECT_TILES_OFF)
        e0, g0 = ps.compute_gradient(ans, obs, x, energy=True)
        lq._lib.call("lq_set_expect_tiles", abi.EXPECT_TILES_ON)
        for _ in range(3):  # eager, captured, replayed
            e1, g1 = ps.compute_gradient(ans, obs, x, energy=True)
            assert abs(e1 - e0) <= 1e-12
            assert np.max(np.abs(np.array(g1) - np.array(g0))) <= 1e-12
    finally:
        modes.restore_defaults()
