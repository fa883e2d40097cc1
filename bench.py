#!/usr/bin/env python3
"""Benchmark: random brickwork circuit, 30 qubits, depth 20, complex128.

A step is one execution of the whole circuit (600 single-qubit rotations +
290 CNOTs) on a 2^30-amplitude state resident in HBM.  `value` is the gate
throughput in GB/s: the algorithmic bytes of every gate (RX/RY/RZ read+write
all amplitudes, 2*c*N; CNOT reads+writes the control=1 half, c*N) summed over
the circuit and divided by the wall time of the timed steps.

The line also carries the roofline of the dominant kernel (the single-qubit
pair kernel, timed live with HIP events on the library stream) and a CPU
baseline: the oracle's single-threaded gate loops (the reference runs its gate
loops on one thread) on a bounded sample of the same 30-qubit workload.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def traffic_from_profiles(mode):
    """HBM bytes per launch of the dominant kernel, measured with rocprofv3
    --pmc (FETCH_SIZE and WRITE_SIZE runs, scaled by the same counters on a
    copy kernel of known size) and recorded in profiles/; None if absent."""
    name = {"per_gate": "traffic_k_pair_rx_30q_c128.txt", "fused": "traffic_lq_pass_30q_c128.txt"}[mode]
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":", 1)[1]))
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=20)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--dtype", choices=["c128", "c64"], default="c128")
    ap.add_argument("--mode", choices=["per_gate", "fused"], default="per_gate")
    ap.add_argument("--cpu-qubits", type=int, default=30)
    ap.add_argument("--cpu-gates", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the gradient / uneven-batch workloads")
    return ap.parse_args()


def algorithmic_bytes(descs, n, c):
    from logosq_amd import abi
    N = 1 << n
    total = 0
    for g in descs:
        if g.kind in (abi.RX, abi.RY, abi.RZ, abi.U3, abi.H, abi.X, abi.Y, abi.MATRIX1, abi.RZZ):
            total += 2 * c * N
        elif g.kind in (abi.CNOT, abi.SWAP, abi.CRX, abi.CRY, abi.CRZ, abi.PHASE, abi.Z, abi.S, abi.T):
            total += c * N
        elif g.kind in (abi.CZ, abi.CPHASE, abi.CP, abi.TOFFOLI):
            total += c * N // 2
        else:
            total += 2 * c * N
    return total


def line_bytes(g, n, c, line=128):
    """HBM bytes a gate moves when every 128-byte line holding an amplitude
    it changes is read and written whole.  A fixed bit (a control, the
    1-half of a phase gate) whose runs of equal value are shorter than a
    line no longer halves the lines touched; neither does a SWAP bit."""
    from logosq_amd import abi
    N = 1 << n
    b = [n - 1 - q for q in list(g.qubits)[:abi.ARITY[g.kind]]]
    fixed = {abi.CNOT: b[:1], abi.CRX: b[:1], abi.CRY: b[:1], abi.CRZ: b[:1], abi.CZ: b, abi.CPHASE: b, abi.CP: b,
             abi.PHASE: b, abi.Z: b, abi.S: b, abi.T: b, abi.TOFFOLI: b[:2]}.get(g.kind, [])
    frac = 1.0
    for p in fixed:
        if (c << p) >= line:
            frac /= 2
    if g.kind == abi.SWAP and all((c << p) >= line for p in b):
        frac /= 2
    return int(2 * c * N * frac)


def gate_kind_sweep(lq, st, n, c, detail=False):
    """Per-gate kernels by gate kind on an n-qubit state: every placement of
    the kind (single-qubit gates on each of the n qubits; two-qubit gates on
    the n-1 neighbour pairs; Toffoli on the n-2 neighbour triples), each
    launch timed with HIP events on the library stream.  Rate = algorithmic
    bytes of the launches (the amplitudes the gate must read and write) over
    their summed time; the slowest placement is reported beside it, and
    the rate against line-granular bytes (line_bytes) as well."""
    from logosq_amd import abi
    kinds = [("RX", abi.RX, 1, (0.3,)), ("RY", abi.RY, 1, (0.3,)), ("RZ", abi.RZ, 1, (0.3,)), ("H", abi.H, 1, ()),
             ("X", abi.X, 1, ()), ("Y", abi.Y, 1, ()), ("Phase", abi.PHASE, 1, (0.4,)),
             ("CNOT", abi.CNOT, 2, ()), ("CZ", abi.CZ, 2, ()), ("CP", abi.CP, 2, (0.5,)),
             ("SWAP", abi.SWAP, 2, ()), ("CRX", abi.CRX, 2, (0.3,)), ("CRZ", abi.CRZ, 2, (0.3,)),
             ("RXX", abi.RXX, 2, (0.3,)), ("RZZ", abi.RZZ, 2, (0.3,)), ("Toffoli", abi.TOFFOLI, 3, ())]
    out = {}
    for name, kind, ar, th in kinds:
        # RXX as (q+1, q): the orientation in which the reference applies
        # the full 4x4 block (src/gates/parameterized.rs:563-575)
        gates = [abi.make_gate(kind, ([q + 1, q] if kind == abi.RXX else list(range(q, q + ar))), th)
                 for q in range(n - ar + 1)]
        tot_b, tot_l, tot_s, worst, rates = 0, 0, 0.0, None, []
        warm = abi.gate_array(gates[:1])
        lq._lib.call("lq_execute", st.handle, warm, 1, abi.EXEC_PER_GATE)  # untimed first launch
        for g in gates:
            arr = abi.gate_array([g])
            s = event_time(lq, lambda: lq._lib.call("lq_execute", st.handle, arr, 1, abi.EXEC_PER_GATE))
            b = algorithmic_bytes([g], n, c)
            tot_b += b
            tot_l += line_bytes(g, n, c)
            tot_s += s
            r = b / s / 1e9
            rates.append(round(r, 1))
            if worst is None or r < worst[0]:
                worst = (r, list(g.qubits)[:ar])
        gbs = tot_b / tot_s / 1e9
        lgbs = tot_l / tot_s / 1e9
        out[name] = {"GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "line_GBps": round(lgbs, 1), "line_frac": round(lgbs / HBM_PEAK_GBS, 4),
                     "bytes_per_launch": algorithmic_bytes(gates[:1], n, c), "launches": len(gates),
                     "slowest_GBps": round(worst[0], 1), "slowest_qubits": worst[1]}
        if detail:
            out[name]["GBps_by_first_qubit"] = rates
    return out


def time_kernels(lq, st, gates, reps=1):
    """GPU time (s) of `reps` per-gate executions of `gates`, from HIP events
    recorded on the library stream around the launches."""
    from logosq_amd import abi
    arr = abi.gate_array(gates)
    lq._lib.call("lq_synchronize")
    lq._lib.call("lq_event_record", 0)
    for _ in range(reps):
        lq._lib.call("lq_execute", st.handle, arr, len(gates), abi.EXEC_PER_GATE)
    lq._lib.call("lq_event_record", 1)
    ms = ctypes.c_float()
    lq._lib.call("lq_event_elapsed_ms", 0, 1, ctypes.byref(ms))
    return ms.value / 1e3


def num_passes(lq, n, dtype, descs):
    from logosq_amd import abi
    order = (ctypes.c_int * len(descs))()
    npass = ctypes.c_int()
    lq._lib.call("lq_plan_fused", n, dtype, abi.gate_array(descs), len(descs), 0, 0, order, None,
                 ctypes.byref(npass))
    return npass.value


def ps_gradient_bench(lq, reps=3, cpu=True):
    """All 2P+1 parameter-shift circuits of a 16-qubit, L=6 ring ansatz
    (P=192) with the 48-term XYZ Heisenberg ring: batched (one launch per
    pass for all 385 circuits) vs the plain sequential GPU method."""
    import numpy as np
    from logosq_amd import abi, algorithms
    tmpl, params, codes, coeffs, j = algorithms.ps_gradient_workload(16, 6, 3)
    arr = abi.gate_array(tmpl)
    P = len(params)
    out = {}
    for name, mode, dtype in (("batched_c128", abi.GRAD_BATCHED, 0), ("batched_c64", abi.GRAD_BATCHED, 1),
                              ("sequential_c128", abi.GRAD_SEQUENTIAL, 0)):
        g = np.zeros(P)
        e = ctypes.c_double()
        ts = []
        for _ in range(reps if mode == abi.GRAD_BATCHED else 1):
            t0 = time.perf_counter()
            lq._lib.call("lq_param_shift_gradient", 16, dtype, arr, len(tmpl),
                         params.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), P, math.pi / 2,
                         codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                         coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(coeffs), 0, mode,
                         ctypes.byref(e), g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            ts.append(time.perf_counter() - t0)
        out[name] = {"s_per_gradient": min(ts), "energy": e.value, "grad": g.copy()}
    ref = out["sequential_c128"]["grad"]
    res = {"workload": "PS gradient, 16 qubits, L=6 RY/RZ+CZ ring, P=192, XYZ ring (48 terms), 385 circuits"}
    for k, v in out.items():
        res[k] = {"ms_per_gradient": round(v["s_per_gradient"] * 1e3, 3),
                  "max_abs_grad_diff_vs_sequential_c128": float(np.max(np.abs(v["grad"] - ref)))}
    res["speedup_batched_vs_sequential"] = round(out["sequential_c128"]["s_per_gradient"] /
                                                 out["batched_c128"]["s_per_gradient"], 2)
    if cpu:
        import oracle
        k = 4  # gradient components timed on the CPU oracle (2k+1 circuits), scaled to 2P+1
        t0 = time.perf_counter()
        e0 = oracle.energy(16, tmpl, params, codes, coeffs)
        for i in range(k):
            pp = params.copy()
            pp[i] += math.pi / 2
            oracle.energy(16, tmpl, pp, codes, coeffs)
            pp[i] -= math.pi
            oracle.energy(16, tmpl, pp, codes, coeffs)
        dt = time.perf_counter() - t0
        res["cpu_oracle_s_per_gradient_extrapolated"] = round(dt / (2 * k + 1) * (2 * P + 1), 2)
        res["cpu_sample"] = "%d of 385 circuits timed, single thread, scaled" % (2 * k + 1)
        res["energy_vs_oracle"] = abs(e0 - out["sequential_c128"]["energy"])
    return res


def uneven_batch_bench(lq, reps=2):
    """512 circuits, n ~ U{10..22}, L=4 ring ansatz, random 32-term Pauli sums:
    complex64 and complex128 batches; per-circuit error of c64 vs c128."""
    import numpy as np
    from logosq_amd import algorithms
    items = algorithms.uneven_batch_workload(512, 10, 22, 4, 32, 4)
    res = {"workload": "512 circuits, n~U{10..22}, L=4 RY/RZ+CZ ring, 32-term random Pauli sums"}
    energies = {}
    for name, dtype in (("c64", 1), ("c128", 0)):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            energies[name] = lq.batch_expectations(items, dtype)
            ts.append(time.perf_counter() - t0)
        res[name] = {"s_per_batch": round(min(ts), 4), "circuits_per_s": round(len(items) / min(ts), 1)}
    d = np.abs(energies["c64"] - energies["c128"])
    res["c64_vs_c128_abs_err"] = {"max": float(d.max()), "mean": float(d.mean())}
    return res


def warmstart_bench(lq, count=4, tile=None):
    """Learned warm-start VQE at 20 qubits: XYZ rings with J ~ U[-1,1]^3,
    the L=8 RY/RZ + CZ-ring ansatz (P=320), network 3-256-256-320 trained on
    the GPU (bf16 MFMA, f32 accumulate) on the committed converged solutions
    of 4096 training couplings.  Time to reach within 1e-4 of the converged
    energy: the plain GPU method (theta0, sequential parameter shift with
    per-gate kernels, as VQE::run with ParameterShift) against the warm-
    started batched method (network proposal with plain fallback, all 2P+1
    shifted circuits batched).  Iteration counts come from the adjoint-
    gradient trajectories, equal to the parameter-shift ones to rounding
    (checked here on the first iterations); per-iteration wall times of both
    methods are measured here."""
    import numpy as np
    from logosq_amd import warmstart
    path = os.path.join(ROOT, "logosq_amd", "data", "warmstart_xyz20_L8.npz")
    evp = os.path.join(ROOT, "logosq_amd", "data", "warmstart_xyz20_L8_eval.json")
    if not os.path.exists(path):
        return {"skipped": "training solutions not present (python -m logosq_amd.warmstart generate)"}
    J, X, E, it, solved, theta0, meta = warmstart.load_data(path)
    if not solved.all():
        return {"skipped": "training solutions incomplete: %d of %d" % (solved.sum(), len(J))}
    ev = None
    if os.path.exists(evp):
        with open(evp) as f:
            ev = json.load(f)
    n, layers = meta["num_qubits"], meta["layers"]
    ans = lq.ry_rz_cz_ring_ansatz(n, layers)
    lr_p = ev["lr_plain"] if ev else meta["learning_rate"]
    lr_w = ev["lr_warm"] if ev else meta["learning_rate"]
    cap, tol = (ev["max_iterations"], ev["tolerance"]) if ev and "max_iterations" in ev else (1500, 1e-10)
    t0 = time.perf_counter()
    # the training recipe the evaluation used (one stage, or the default schedule)
    ws, hist = warmstart.WarmStart.train(J, X, theta0, epochs=ev.get("epochs") if ev else None,
                                         learning_rate=ev.get("lr") if ev else None)
    t_train = time.perf_counter() - t0
    Jte = warmstart.couplings(meta["train"] + meta["held_out"], meta["seed"])[meta["train"]:][:count]
    # the recorded evaluation ran the generic tile kernel at the default tile
    # geometry, and so does its live re-check (equal iteration counts show
    # the record reproduces; the runs are chaotic, so any other rounding
    # moves them)
    lq.jit.set_jit(lq.jit.OFF)
    lq._lib.call("lq_set_fused_config", 0, 0, 0)
    try:
        rows = warmstart.evaluate(ans, ws, Jte, theta0, lr_p, lr_w, cap, tol)
    finally:
        lq.jit.set_jit(lq.jit.REPEAT)
        if tile:
            lq._lib.call("lq_set_fused_config", *tile)
    # per-iteration wall time of each method at the first held-out coupling
    obs = warmstart.hamiltonians(n, Jte[:1])[0]
    seq, bat, adj = lq.ParameterShift(batched=False), lq.ParameterShift(batched=True), lq.AdjointGradient()

    def wall(fn, reps):
        fn()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return min(ts)
    t_seq = wall(lambda: seq.compute_gradient(ans, obs, theta0, energy=True), 1)
    t_seqf = wall(lambda: sequential_fused_ps(lq, ans, obs, theta0), 1)
    g_sf = sequential_fused_ps(lq, ans, obs, theta0)[1]
    g_bat = np.array(bat.compute_gradient(ans, obs, theta0))
    t_bat = wall(lambda: bat.compute_gradient(ans, obs, theta0, energy=True), 2)
    lq.jit.set_jit(lq.jit.OFF)
    t_bat_generic = wall(lambda: bat.compute_gradient(ans, obs, theta0, energy=True), 2)
    lq.jit.set_jit(lq.jit.REPEAT)
    t_adj = wall(lambda: adj.compute_gradient(ans, obs, theta0, energy=True), 3)
    t_start = wall(lambda: warmstart.choose_starts(ws, ans, Jte[:1]), 2)
    # the batched parameter-shift VQE follows the adjoint trajectory
    start = warmstart.choose_starts(ws, ans, Jte[:1])[0][0]
    r_ps = lq.VQE(ans, obs, bat, lq.Adam(lr_w, 3)).run(start)
    r_adj = lq.VQE(ans, obs, adj, lq.Adam(lr_w, 3)).run(start)
    traj = float(np.max(np.abs(np.array(r_ps.convergence_history) - r_adj.convergence_history)))

    def times(rs, kp, kw, t_plain=t_seq, t_warm=t_bat):
        tp = [(r[kp] or cap) * t_plain for r in rs]
        tw = [t_start + (r[kw] or cap) * t_warm for r in rs]
        return tp, tw

    def summary(rs):
        out = {}
        # dE < 1e-4 against the plain method's converged f64 energy (primary),
        # and against the lower of the two methods' converged energies
        for name, kp, kw in (("vs_plain_converged_energy", "iters_plain_to_plain", "iters_warm_to_plain"),
                             ("vs_best_converged_energy", "iters_plain", "iters_warm")):
            tp, tw = times(rs, kp, kw)
            out[name] = {"reached_plain": sum(r[kp] is not None for r in rs),
                         "reached_warm": sum(r[kw] is not None for r in rs),
                         "median_iterations_plain": float(np.median([r[kp] or cap for r in rs])),
                         "median_iterations_warm": float(np.median([r[kw] or cap for r in rs])),
                         "median_time_to_dE_s_plain": round(float(np.median(tp)), 2),
                         "median_time_to_dE_s_warm_batched": round(float(np.median(tw)), 2),
                         "speedup_median": round(float(np.median(tp) / np.median(tw)), 2),
                         "speedup_total": round(float(np.sum(tp) / np.sum(tw)), 2)}
            # the same iterations against the strongest one-circuit-at-a-time
            # plain method (fused execution), and with the adjoint gradient
            tpf, _ = times(rs, kp, kw, t_plain=t_seqf)
            _, twa = times(rs, kp, kw, t_warm=t_adj)
            out[name]["speedup_median_vs_plain_fused_sequential"] = round(float(np.median(tpf) / np.median(tw)), 2)
            out[name]["speedup_median_warm_adjoint_vs_plain"] = round(float(np.median(tp) / np.median(twa)), 1)
        out["warm_final_energy_lower"] = sum(r["e_warm_final"] < r["e_plain_final"] - 1e-4 for r in rs)
        out["proposal_used"] = sum(r["used_proposal"] for r in rs)
        return out

    out = {"workload": "20q XYZ ring, L=8 RY/RZ+CZ (P=320), 4096 training solutions; %d held-out couplings run "
                       "live, iteration cap %d" % (len(rows), cap),
           "train_s": round(t_train, 2), "train_loss_first_last": [float(hist[0]), float(hist[-1])],
           "lr_plain": lr_p, "lr_warm": lr_w,
           "s_per_iteration": {"plain_sequential_ps": round(t_seq, 4),
                               "plain_sequential_ps_fused_execution": round(t_seqf, 4),
                               "batched_ps": round(t_bat, 4),
                               "batched_ps_generic_tile_kernel": round(t_bat_generic, 4),
                               "adjoint": round(t_adj, 5), "warm_start_proposal": round(t_start, 4)},
           "trajectory_batched_ps_vs_adjoint_max_abs": traj,
           "gradient_fused_sequential_vs_batched_max_abs": float(np.max(np.abs(g_sf - g_bat))),
           "live": summary(rows)}
    if ev:
        # the live couplings are the first held-out ones: same iterations
        rec = {tuple(r["J"]): r for r in ev["rows"]}
        same = [rec[tuple(r["J"])]["iters_warm_to_plain"] == r["iters_warm_to_plain"] and
                rec[tuple(r["J"])]["iters_plain_to_plain"] == r["iters_plain_to_plain"]
                for r in rows if tuple(r["J"]) in rec]
        out["live_iterations_equal_recorded"] = "%d of %d" % (sum(same), len(same))
        out["held_out_%d" % len(ev["rows"])] = summary(ev["rows"])
        out["held_out_%d" % len(ev["rows"])]["source"] = ("logosq_amd/data/warmstart_xyz20_L8_eval.json (iterations), "
                                                        "per-iteration times measured here")
    return out


def sequential_fused_ps(lq, ans, obs, theta, shift=math.pi / 2):
    """One parameter-shift gradient with the 2P+1 circuits run one after
    another, each with fused tile-pass execution (lq_execute FUSED) and one
    expectation: the strongest one-circuit-at-a-time GPU method.  The bound
    gate array is built once; only the shifted angle changes per circuit."""
    import numpy as np
    from logosq_amd import abi, algorithms
    tmpl = ans.template()
    n = ans.num_qubits()
    arr = abi.gate_array(algorithms.bind(tmpl, theta))
    slot = {}
    for i, g in enumerate(tmpl):
        for k in range(3):
            if g.param[k] >= 0:
                slot[g.param[k]] = (i, k)
    codes, coeffs = obs.codes()
    st = lq.State(n)
    e = ctypes.c_double()
    cp = codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    fp = coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def energy():
        lq._lib.call("lq_state_set_basis", st.handle, 0)
        lq._lib.call("lq_execute", st.handle, arr, len(tmpl), abi.EXEC_FUSED)
        lq._lib.call("lq_expect_pauli_sum", st.handle, cp, fp, len(coeffs), obs.convention, ctypes.byref(e))
        return e.value
    e0 = energy()
    grad = np.zeros(len(theta))
    for p, (i, k) in slot.items():
        arr[i].theta[k] = theta[p] + shift
        ep = energy()
        arr[i].theta[k] = theta[p] - shift
        em = energy()
        arr[i].theta[k] = theta[p]
        grad[p] = (ep - em) / 2.0
    return e0, grad


def cpu_baseline(nq, ngates, seed, depth):
    """The oracle's restated gate loops, single thread, on the first gates of
    the same circuit at nq qubits."""
    import numpy as np
    import oracle
    from logosq_amd import algorithms, abi
    lib = oracle.lib()
    descs = algorithms.brickwork_circuit(nq, depth, seed).descriptors()
    layer0 = descs[:nq] + descs[nq:nq + nq // 2]
    nrot = (2 * ngates) // 3
    sample = layer0[:min(nrot, nq)] + layer0[nq:nq + ngates - min(nrot, nq)]
    v = np.zeros(1 << nq, dtype=np.complex128)
    v[:] = 0  # touch every page before timing
    v[0] = 1
    p = v.ctypes.data_as(ctypes.c_void_p)
    total_bytes = algorithmic_bytes(sample, nq, 16)
    t0 = time.perf_counter()
    for g in sample:
        lib.or_apply_gate(p, nq, ctypes.byref(g))
    dt = time.perf_counter() - t0
    names = {abi.RX: "RX", abi.RY: "RY", abi.RZ: "RZ", abi.CNOT: "CNOT"}
    return {"value": total_bytes / dt / 1e9, "unit": "GB/s", "cores": 1, "kind": "port", "host": host_info(),
            "sample": "%d gates of layer 0 (%s) at %d qubits c128, oracle gate loops single-threaded as the "
                      "reference's, %.1f s" % (len(sample), ",".join(names.get(g.kind, str(g.kind)) for g in sample),
                                              nq, dt)}


def host_info():
    """The GPU box's host: logical CPUs, the CPUs this process may use, and
    the CPU model (lscpu, else /proc/cpuinfo)."""
    import subprocess
    info = {"nproc": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0)), "cpu_model": None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if not info["cpu_model"] and os.path.exists("/proc/cpuinfo"):
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
                break
    return info


def copy_peak(lq, st, n, dtype, c, reps=3):
    """Measured device copy rate: lq_state_copy of the whole 2^n-amplitude
    state into a second state (the library's streaming copy kernel), read +
    write bytes over the fastest of `reps` HIP-event times."""
    dst = lq.State(n, dtype)
    fn = lambda: lq._lib.call("lq_state_copy", dst.handle, st.handle)
    fn()
    t = min(event_time(lq, fn) for _ in range(reps))
    dst = None
    return 2 * c * (1 << n) / t / 1e9


def run_circuit(lq, st, arr, count, mode, steps, warmup, label):
    from logosq_amd import abi
    for _ in range(warmup):
        lq._lib.call("lq_state_set_basis", st.handle, 0)
        lq._lib.call("lq_execute", st.handle, arr, count, mode)
    lq._lib.call("lq_synchronize")
    times = []
    for _ in range(steps):
        lq._lib.call("lq_state_set_basis", st.handle, 0)
        lq._lib.call("lq_synchronize")
        t0 = time.perf_counter()
        lq._lib.call("lq_execute", st.handle, arr, count, mode)
        lq._lib.call("lq_synchronize")
        times.append(time.perf_counter() - t0)
        print("bench: %s step %.3f s" % (label, times[-1]), file=sys.stderr, flush=True)
    return times


def expectation_bench(lq, st, n, c, reps=3):
    """Pauli-sum expectation on the circuit's output state: the periodic XYZ
    ring, 3n terms with J = (0.7, -0.4, 0.9).  The XX and YY terms of a bond
    share one flip mask and every ZZ term is diagonal: n + 1 groups.  The
    groups are evaluated from LDS tiles by kernels specialised at run time
    (compiled before timing), a few layouts each reading the state once, the
    diagonal group inside the first; `reads` counts the state reads (layouts
    + groups read alone + diagonal groups read alone), `bytes` = reads * c *
    N.  HIP-event time of the whole lq_expect_pauli_sum call (fastest of
    `reps`); beside it the generic tile kernels and one read per group."""
    import numpy as np
    codes, coeffs = lq.algorithms.heisenberg_codes(n, 0.7, -0.4, 0.9)
    groups = len({tuple(((row == 1) | (row == 2)).tolist()) for row in codes})
    e = ctypes.c_double()
    cp = codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    fp = coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    dtype = 0 if c == 16 else 1
    kernels = ctypes.c_int()
    lq._lib.call("lq_expect_jit_compile", n, dtype, cp, fp, len(coeffs), 0, ctypes.byref(kernels))
    lay, pair, diag = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lq._lib.call("lq_expect_plan", n, dtype, cp, fp, len(coeffs), 0, ctypes.byref(lay), ctypes.byref(pair),
                 ctypes.byref(diag))
    reads = lay.value + pair.value + diag.value
    fn = lambda: lq._lib.call("lq_expect_pauli_sum", st.handle, cp, fp, len(coeffs), 0, ctypes.byref(e))
    fn()
    t = min(event_time(lq, fn) for _ in range(reps))
    e_spec = e.value
    lq._lib.call("lq_set_jit", 0)  # the generic tile kernels, for comparison
    try:
        fn()
        t_tiles = min(event_time(lq, fn) for _ in range(reps))
        e_tiles = e.value
    finally:
        lq._lib.call("lq_set_jit", 1)
    lq._lib.call("lq_set_expect_tiles", 0)  # one read per group, for comparison
    try:
        fn()
        t_groups = min(event_time(lq, fn) for _ in range(reps))
    finally:
        lq._lib.call("lq_set_expect_tiles", 1)
    gbs = reads * c * (1 << n) / t / 1e9
    return {"workload": "periodic XYZ ring, %d terms, %d groups: %d tile layouts (specialised kernels: %d), %d "
                        "flip groups read alone, %d diagonal groups read alone"
                        % (len(coeffs), groups, lay.value, kernels.value, pair.value, diag.value),
            "ms": round(t * 1e3, 3), "reads": reads, "bytes": reads * c * (1 << n), "GBps": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4), "ms_generic_tile_kernels": round(t_tiles * 1e3, 3),
            "ms_one_read_per_group": round(t_groups * 1e3, 3), "energy": e_spec,
            "energy_diff_generic": abs(e_spec - e_tiles), "energy_diff_one_read_per_group": abs(e_spec - e.value)}


def c64_error(lq, n, depth, seed):
    """The brickwork circuit at n qubits in complex64 (per-gate and fused)
    against complex128: max |amplitude difference| and the f64-accumulated
    norm of the complex64 result."""
    import numpy as np
    from logosq_amd import abi
    descs = lq.algorithms.brickwork_circuit(n, depth, seed).descriptors()
    arr = abi.gate_array(descs)
    s128 = lq.State(n, lq.C128)
    lq._lib.call("lq_execute", s128.handle, arr, len(descs), abi.EXEC_PER_GATE)
    ref = s128.vector()
    s128 = None
    out = {"qubits": n}
    for name, mode in (("per_gate", abi.EXEC_PER_GATE), ("fused", abi.EXEC_FUSED)):
        s = lq.State(n, lq.C64)
        lq._lib.call("lq_execute", s.handle, arr, len(descs), mode)
        out[name + "_max_abs_err"] = float(np.max(np.abs(s.vector() - ref)))
        out[name + "_norm_sqr"] = s.norm_sqr()
    return out


def event_time(lq, fn):
    """GPU time (s) of the launches fn() queues, from HIP events recorded on
    the library stream around them."""
    lq._lib.call("lq_synchronize")
    lq._lib.call("lq_event_record", 2)
    fn()
    lq._lib.call("lq_event_record", 3)
    ms = ctypes.c_float()
    lq._lib.call("lq_event_elapsed_ms", 2, 3, ctypes.byref(ms))
    return ms.value / 1e3


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (--gpus %d requested)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    import logosq_amd as lq
    from logosq_amd import abi
    lq.load()
    dtype = lq.C128 if a.dtype == "c128" else lq.C64
    c = 16 if dtype == lq.C128 else 8
    mode = abi.EXEC_PER_GATE if a.mode == "per_gate" else abi.EXEC_FUSED
    tune = None
    if not a.no_tune:
        buf = ctypes.create_string_buffer(8192)
        lq._lib.call("lq_autotune", a.qubits, dtype, buf, 8192)
        tune = json.loads(buf.value.decode())
        print("bench: autotune %s %s" % (tune["gate_choice"], tune["tile_choice"]), file=sys.stderr, flush=True)
    circ = lq.algorithms.brickwork_circuit(a.qubits, a.depth, a.seed)
    descs = circ.descriptors()
    arr = abi.gate_array(descs)
    # the fused engine's pass kernels, specialised per pass structure and
    # compiled before any timed run
    t = time.perf_counter()
    jit_kernels = lq.jit.compile_plan(a.qubits, descs, dtype)
    jit_compile_s = time.perf_counter() - t
    jit0 = lq.jit.jit_stats()
    st = lq.State(a.qubits, dtype)
    N = 1 << a.qubits
    print("bench: %d qubits %s depth %d, %d gates, mode %s" % (a.qubits, a.dtype, a.depth, len(descs), a.mode),
          file=sys.stderr, flush=True)
    times = run_circuit(lq, st, arr, len(descs), mode, a.steps, a.warmup, a.mode)
    norm2 = st.norm_sqr()
    copy_gbs = copy_peak(lq, st, a.qubits, dtype, c)
    print("bench: device copy %.1f GB/s" % copy_gbs, file=sys.stderr, flush=True)
    expect = expectation_bench(lq, st, a.qubits, c)
    print("bench: expectation %.1f GB/s" % expect["GBps"], file=sys.stderr, flush=True)
    step_s = sum(times) / len(times)
    bytes_gates = algorithmic_bytes(descs, a.qubits, c)
    passes = num_passes(lq, a.qubits, dtype, descs)
    if mode == abi.EXEC_FUSED:
        bytes_moved = passes * 2 * c * N
    else:
        bytes_moved = bytes_gates
    value = bytes_moved / step_s / 1e9
    # per-gate kernels: RX on every target (pair kernel) and the circuit's CNOTs
    rx = [abi.make_gate(abi.RX, [q], [0.1 + q]) for q in range(a.qubits)]
    cn = [g for g in descs if g.kind == abi.CNOT][: a.qubits // 2]
    rx_arr, cn_arr = abi.gate_array(rx), abi.gate_array(cn)
    t_rx = event_time(lq, lambda: lq._lib.call("lq_execute", st.handle, rx_arr, len(rx), abi.EXEC_PER_GATE))
    t_cn = event_time(lq, lambda: lq._lib.call("lq_execute", st.handle, cn_arr, len(cn), abi.EXEC_PER_GATE))
    rx_gbs = 2 * c * N * len(rx) / t_rx / 1e9
    cn_gbs = c * N * len(cn) / t_cn / 1e9
    by_kind = None
    if not a.no_extra:
        print("bench: per-gate kernels by kind", file=sys.stderr, flush=True)
        by_kind = gate_kind_sweep(lq, st, a.qubits, c)
    # the other execution mode of the same circuit
    other = abi.EXEC_FUSED if mode == abi.EXEC_PER_GATE else abi.EXEC_PER_GATE
    oname = "fused" if other == abi.EXEC_FUSED else "per_gate"
    t_other = run_circuit(lq, st, arr, len(descs), other, 1, 1 if other == abi.EXEC_FUSED else 0, oname)[0]
    t_pass = event_time(lq, lambda: (lq._lib.call("lq_state_set_basis", st.handle, 0),
                                     lq._lib.call("lq_execute", st.handle, arr, len(descs), abi.EXEC_FUSED)))
    t_fill = event_time(lq, lambda: lq._lib.call("lq_state_set_basis", st.handle, 0))
    pass_s = max(t_pass - t_fill, 1e-9) / passes
    jit_launches = lq.jit.jit_stats()["launches"] - jit0["launches"]
    fused_s = t_other if other == abi.EXEC_FUSED else step_s
    per_gate_s = step_s if mode == abi.EXEC_PER_GATE else t_other
    if mode == abi.EXEC_PER_GATE:
        kname = "k_pair_hi / k_pair_lo (RX on each of the %d targets, per-gate path)" % a.qubits
        bpl, avg_s = 2 * c * N, t_rx / len(rx)
    else:
        kname = ("lq_pass (fused LDS tile pass specialised per pass structure), %d passes per circuit" % passes
                 if jit_launches else "k_tileg (fused LDS tile pass), %d passes per circuit" % passes)
        bpl, avg_s = 2 * c * N, pass_s
    achieved = bpl / avg_s / 1e9
    fused = {"ms_per_circuit": round(fused_s * 1e3, 3), "passes": passes,
             "speedup_vs_per_gate": round(per_gate_s / fused_s, 2),
             "effective_gate_GBps": round(bytes_gates / fused_s / 1e9, 1),
             "pass_kernel_GBps": round(2 * c * N / pass_s / 1e9, 1),
             "pass_kernel_hbm_frac": round(2 * c * N / pass_s / 1e9 / HBM_PEAK_GBS, 4),
             "pass_kernel": "lq_pass (specialised)" if jit_launches else "k_tileg (generic)",
             "specialised_kernels": jit_kernels, "specialised_compile_s": round(jit_compile_s, 2),
             "specialised_launches": jit_launches}
    st = None
    c64 = None
    if dtype == lq.C128 and not a.no_extra:
        # the same per-gate kernels on a complex64 state
        s64 = lq.State(a.qubits, lq.C64)
        t64 = event_time(lq, lambda: lq._lib.call("lq_execute", s64.handle, rx_arr, len(rx), abi.EXEC_PER_GATE))
        t64c = event_time(lq, lambda: lq._lib.call("lq_execute", s64.handle, cn_arr, len(cn), abi.EXEC_PER_GATE))
        c64 = {"rx_pair_GBps": round(2 * 8 * N * len(rx) / t64 / 1e9, 1),
               "cnot_GBps": round(8 * N * len(cn) / t64c / 1e9, 1)}
        s64 = None
        c64["error_vs_c128"] = c64_error(lq, min(a.qubits, 24), a.depth, a.seed)
    extra = {}
    if not a.no_extra:
        print("bench: gradient workload", file=sys.stderr, flush=True)
        extra["ps_gradient_16q"] = ps_gradient_bench(lq, cpu=not a.no_cpu)
        print("bench: uneven batch workload", file=sys.stderr, flush=True)
        extra["uneven_batch"] = uneven_batch_bench(lq)
        print("bench: warm-start VQE workload", file=sys.stderr, flush=True)
        tile = None
        if tune:
            tc = tune["tile_choice"]
            tile = ((tc["tile_bits"], 0, tc["low_bits"]) if dtype == lq.C128 else (0, tc["tile_bits"], tc["low_bits"]))
        extra["warmstart_vqe_20q"] = warmstart_bench(lq, tile=tile)
    cpu = None
    if not a.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(a.cpu_qubits, a.cpu_gates, a.seed, a.depth)
    out = {
        "metric": "gate HBM GB/s at 30q c128; VQE time-to-dE<1e-4, warm-start+batched vs plain",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": a.dtype,
        "data": "synthetic",
        "config": {"workload": "random brickwork circuit, %d qubits, depth %d, %s, seed %d, %s execution"
                               % (a.qubits, a.depth, a.dtype, a.seed, a.mode.replace("_", "-")),
                   "qubits": a.qubits, "depth": a.depth, "gates": len(descs), "autotune": tune,
                   "hbm_bytes_per_step": bytes_moved, "gate_algorithmic_bytes_per_step": bytes_gates,
                   "final_norm_sqr": norm2,
                   "per_gate_kernels": {"rx_pair_GBps": round(rx_gbs, 1), "rx_frac": round(rx_gbs / HBM_PEAK_GBS, 4),
                                        "cnot_GBps": round(cn_gbs, 1), "cnot_frac": round(cn_gbs / HBM_PEAK_GBS, 4),
                                        "c64": c64, "by_kind": by_kind},
                   "fused_execution": fused,
                   "expectation": expect},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(a.mode),
                     "kernel": "%s, %d B per launch, avg %.3f ms" % (kname, bpl, avg_s * 1e3),
                     "measured_peak": round(copy_gbs, 1),
                     "measured_peak_kernel": "lq_state_copy of the %d-qubit state (k_copy16: one 16-byte unit per "
                                             "thread, nontemporal, read + write bytes)" % a.qubits,
                     "frac_vs_measured": round(achieved / copy_gbs, 4)},
        "cpu_baseline": cpu,
        "host": host_info(),
        "workloads": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
