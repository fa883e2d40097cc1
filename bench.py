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
    ap.add_argument("--workload", choices=["brickwork", "uneven", "gradient"], default="brickwork",
                    help="brickwork: the headline run; uneven / gradient: only that workload (one JSON line)")
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


FP64_PEAK_GFLOPS = 78600.0  # MI355X dense FP64 vector (FMA counted as two flops)


def algorithmic_flops(descs, n):
    """Floating-point operations of a circuit's gates at their minimum
    count per amplitude: a real-coefficient rotation (RX, RY: a multiply and
    a multiply-add per output value) 6, a phase gate applied to every
    amplitude (RZ: one complex multiply) 6, a CNOT none (a permutation);
    other kinds are not counted (None)."""
    from logosq_amd import abi
    per = {abi.RX: 6, abi.RY: 6, abi.RZ: 6, abi.CNOT: 0}
    total = 0
    for g in descs:
        if g.kind not in per:
            return None
        total += per[g.kind] * (1 << n)
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


def num_passes(lq, n, dtype, descs, tile_bits=0):
    from logosq_amd import abi
    order = (ctypes.c_int * len(descs))()
    npass = ctypes.c_int()
    lq._lib.call("lq_plan_fused", n, dtype, abi.gate_array(descs), len(descs), tile_bits, 0, order, None,
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
        # batched: the first call runs the generic pass kernel, the second
        # compiles the specialised ones, the third captures the launch
        # sequence; the steady state is the fastest of `reps` calls after that
        calls = reps + 3 if mode == abi.GRAD_BATCHED else 1
        for _ in range(calls):
            t0 = time.perf_counter()
            lq._lib.call("lq_param_shift_gradient", 16, dtype, arr, len(tmpl),
                         params.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), P, math.pi / 2,
                         codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                         coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(coeffs), 0, mode,
                         ctypes.byref(e), g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            ts.append(time.perf_counter() - t0)
        out[name] = {"s_per_gradient": min(ts[-reps:]) if mode == abi.GRAD_BATCHED else ts[0],
                     "first_call_s": ts[0], "energy": e.value, "grad": g.copy()}
    ref = out["sequential_c128"]["grad"]
    res = {"workload": "PS gradient, 16 qubits, L=6 RY/RZ+CZ ring, P=192, XYZ ring (48 terms), 385 circuits"}
    for k, v in out.items():
        res[k] = {"ms_per_gradient": round(v["s_per_gradient"] * 1e3, 3),
                  "max_abs_grad_diff_vs_sequential_c128": float(np.max(np.abs(v["grad"] - ref)))}
        if k.startswith("batched"):
            res[k]["first_call_ms"] = round(v["first_call_s"] * 1e3, 3)
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


def uneven_batch_bench(lq, reps=3):
    """512 circuits, n ~ U{10..22}, L=4 ring ansatz, random 32-term Pauli sums:
    complex64 and complex128 batches (lq_batch_expectations); per-circuit
    error of c64 vs c128.  The arguments are marshalled once (Python ctypes
    arrays: timed apart); the first call of a batch shape runs eagerly and
    the second compiles its specialised pass kernels and captures its
    graphs, so the steady state is the fastest of `reps` calls after two;
    the first call is reported beside it.  `state_GBps`: the bytes of all 512
    states (sum of 2^n amplitudes x c) per second of one call."""
    import numpy as np
    from logosq_amd import algorithms
    from logosq_amd.batch import BatchArgs
    items = algorithms.uneven_batch_workload(512, 10, 22, 4, 32, 4)
    t0 = time.perf_counter()
    args = BatchArgs(items)
    t_marshal = time.perf_counter() - t0
    amps = sum(1 << it[0] for it in items)
    res = {"workload": "512 circuits, n~U{10..22}, L=4 RY/RZ+CZ ring, 32-term random Pauli sums",
           "python_marshalling_s": round(t_marshal, 4), "total_amplitudes": amps}
    energies = {}
    for name, dtype in (("c64", 1), ("c128", 0)):
        c = 16 if dtype == 0 else 8
        t0 = time.perf_counter()
        lq.batch_expectations(args, dtype)
        t_first = time.perf_counter() - t0
        lq.batch_expectations(args, dtype)  # compiles the specialised passes, captures the graphs
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            energies[name] = lq.batch_expectations(args, dtype)
            ts.append(time.perf_counter() - t0)
        best = min(ts)
        res[name] = {"s_per_batch": round(best, 4), "circuits_per_s": round(len(items) / best, 1),
                     "state_GBps": round(amps * c / best / 1e9, 1), "first_call_s": round(t_first, 4)}
    d = np.abs(energies["c64"] - energies["c128"])
    res["c64_vs_c128_abs_err"] = {"max": float(d.max()), "mean": float(d.mean())}
    # state traffic the batch implies, counted as if every access went to
    # HBM (small states are served from the caches): per circuit, its fused
    # plan's passes read and write the state once each (the first pass only
    # writes: it builds |0...0> in LDS), and its Pauli sum reads it once per
    # orbit item (up to four linearly independent flip masks per read)
    by_n = {}
    for n, gates, codes, coeffs in items:
        if n not in by_n:
            # batches plan complex64 states with 2^12-amplitude tiles (the
            # library's batch default), complex128 with the default 2^12
            by_n[n] = {0: num_passes(lq, n, 0, gates), 1: num_passes(lq, n, 1, gates, min(12, n))}
    reads = []
    for n, gates, codes, coeffs in items:
        masks = set()
        for row in codes:
            x = 0
            for q, p in enumerate(row):
                if p in (1, 2):
                    x |= 1 << q
            if x:
                masks.add(x)
        reads.append(max(1, -(-len(masks) // 4)))
    for name, dtype in (("c64", 1), ("c128", 0)):
        c = 16 if dtype == 0 else 8
        traffic = sum(((2 * by_n[it[0]][dtype] - 1) + r) * (1 << it[0]) * c for it, r in zip(items, reads))
        res[name]["modelled_state_traffic_GB"] = round(traffic / 1e9, 2)
        res[name]["modelled_state_traffic_GBps"] = round(traffic / res[name]["s_per_batch"] / 1e9, 1)
    return res


def warmstart_bench(lq, live=2, same_method=2, tile=None):
    """Learned warm-start VQE at 20 qubits: XYZ rings with J ~ U[-1,1]^3,
    the L=8 RY/RZ + CZ-ring ansatz (P=320), network 3-256-256-320 trained on
    the GPU here (bf16 MFMA, f32 accumulate) on the committed converged
    solutions of 4096 training couplings.

    Target: within 1e-4 of E*, the lower final energy of two long runs
    (Adam, adjoint gradient, 5000 iterations, from the plain and from the
    warm start) of each of the first 32 held-out couplings, committed in
    logosq_amd/data/warmstart_xyz20_L8_converged.json.  No run stops on
    Adam's own tests within 5000 iterations, so E* is the best energy these
    runs reach, not a certified minimum.

    * Modelled over all 32 couplings: recorded iterations to the target x
      measured per-iteration time of each method (plain: VQE::run with the
      sequential parameter shift on the per-gate kernels; warm: network
      proposal + VQE::run with the batched parameter shift).
    * Measured end to end, as a check of that model: the warm-started
      batched method on `live` couplings (the first whose recorded warm run
      reached the target within 1000 iterations) -- network proposal, start
      choice (f64 energies), VQE::run with the batched parameter shift (all
      2P+1 circuits per pass launch), stopped at the first energy within the
      target (lq_vqe_run_to_target) -- wall time.
    * The warm start's own gain, measured end to end with the same gradient
      method on both sides (adjoint): plain start against warm start to the
      same target, first `same_method` couplings.
    * The plain GPU method (VQE::run with sequential parameter shift on the
      per-gate kernels, 2P+1 circuits one after another) takes ~1.8 s per
      iteration and thousands of iterations: its time is modelled as the
      recorded iterations x its measured per-iteration time, and labelled so."""
    import numpy as np
    from logosq_amd import warmstart
    path = os.path.join(ROOT, "logosq_amd", "data", "warmstart_xyz20_L8.npz")
    evp = os.path.join(ROOT, "logosq_amd", "data", "warmstart_xyz20_L8_eval.json")
    cvp = os.path.join(ROOT, "logosq_amd", "data", "warmstart_xyz20_L8_converged.json")
    for pth in (path, evp, cvp):
        if not os.path.exists(pth):
            return {"skipped": "%s not present" % os.path.relpath(pth, ROOT)}
    J, X, E, it, solved, theta0, meta = warmstart.load_data(path)
    with open(evp) as f:
        ev = json.load(f)
    with open(cvp) as f:
        cv = json.load(f)
    n, layers = meta["num_qubits"], meta["layers"]
    ans = lq.ry_rz_cz_ring_ansatz(n, layers)
    lr_p, lr_w = cv["lr_plain"], cv["lr_warm"]
    cap = cv["max_iterations"]
    t0 = time.perf_counter()
    ws, hist = warmstart.WarmStart.train(J, X, theta0, epochs=cv.get("epochs"), learning_rate=cv.get("lr"))
    t_train = time.perf_counter() - t0
    rows = cv["rows"]
    Jc = np.array([r["J"] for r in rows])
    if tile:
        lq._lib.call("lq_set_fused_config", *tile)
    obs0 = warmstart.hamiltonians(n, Jc[:1])[0]
    seq, bat, adj = lq.ParameterShift(batched=False), lq.ParameterShift(batched=True), lq.AdjointGradient()

    def wall(fn, reps):
        fn()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return min(ts)
    # per-iteration times (one gradient with its energy) at the plain start
    t_seq = wall(lambda: seq.compute_gradient(ans, obs0, theta0, energy=True), 1)
    t_seqf = wall(lambda: sequential_fused_ps(lq, ans, obs0, theta0), 1)
    g_sf = sequential_fused_ps(lq, ans, obs0, theta0)[1]
    g_bat = np.array(bat.compute_gradient(ans, obs0, theta0))
    t_bat = wall(lambda: bat.compute_gradient(ans, obs0, theta0, energy=True), 2)
    t_adj = wall(lambda: adj.compute_gradient(ans, obs0, theta0, energy=True), 3)
    t_start = wall(lambda: warmstart.choose_starts(ws, ans, Jc[:1]), 2)

    def to_target(j, row, method, lr, warm, max_it):
        """Wall time from the couplings to the first energy within the
        target: (seconds, iterations, reached)."""
        obs = warmstart.hamiltonians(n, [j])[0]
        t = time.perf_counter()
        x0 = warmstart.choose_starts(ws, ans, [j])[0][0] if warm else theta0
        opt = lq.Adam(lr, max_it).with_tolerance(cv["tolerance"])  # the record's stopping tests
        r, reached = lq.VQE(ans, obs, method, opt).run_to_target(x0, row["e_star"] + 1e-4)
        lq._lib.call("lq_synchronize")
        return time.perf_counter() - t, len(r.convergence_history), reached

    # live sample: warm runs take ~50 ms per iteration and thousands of
    # iterations for most couplings, so the end-to-end check of the model
    # runs on the first couplings whose recorded warm run reached the target
    # within 1000 iterations
    pick = [i for i, r in enumerate(rows) if r["iters_warm"] is not None and r["iters_warm"] <= 1000][:live]
    measured = []
    for j, row in zip(Jc[pick], [rows[i] for i in pick]):
        max_it = min(cap, 2 * (row["iters_warm"] or cap) + 200)
        s_, k_, ok = to_target(j, row, bat, lr_w, True, max_it)
        measured.append({"coupling": pick[len(measured)], "J": [round(v, 6) for v in j], "s": round(s_, 3),
                         "iterations": k_, "reached": ok,
                         "modelled_s": round(t_start + (row["iters_warm"] or cap) * t_bat, 2),
                         "recorded_iterations_warm": row["iters_warm"],
                         "plain_modelled_s": round((row["iters_plain"] or cap) * t_seq, 1),
                         "recorded_iterations_plain": row["iters_plain"]})
        print("bench: warm batched VQE to target %.2f s (%d iterations)" % (s_, k_), file=sys.stderr, flush=True)
    own = []
    for j, row in zip(Jc[:same_method], rows[:same_method]):
        sp, kp, okp = to_target(j, row, adj, lr_p, False, cap)
        sw, kw, okw = to_target(j, row, adj, lr_w, True, cap)
        own.append({"plain_s": round(sp, 3), "plain_iterations": kp, "plain_reached": okp,
                    "warm_s": round(sw, 3), "warm_iterations": kw, "warm_reached": okw})
    med = lambda v: float(np.median(v))
    ip = [r["iters_plain"] or cap for r in rows]
    iw = [r["iters_warm"] or cap for r in rows]
    tp_model = [k * t_seq for k in ip]
    tw_model = [t_start + k * t_bat for k in iw]
    out = {"workload": "20q XYZ ring, L=8 RY/RZ+CZ (P=320), 4096 training solutions; E* from runs to convergence "
                       "(%d held-out couplings, both starts, up to %d iterations)" % (len(rows), cap),
           "train_s": round(t_train, 2), "train_loss_first_last": [float(hist[0]), float(hist[-1])],
           "lr_plain": lr_p, "lr_warm": lr_w,
           "learning_rates_from": "validation grid (0.02, 0.05, 0.1) per method, 16 couplings of another stream",
           "s_per_iteration": {"plain_sequential_ps": round(t_seq, 4),
                               "plain_sequential_ps_fused_execution": round(t_seqf, 4),
                               "batched_ps": round(t_bat, 4), "adjoint": round(t_adj, 5),
                               "warm_start_proposal": round(t_start, 4)},
           "gradient_fused_sequential_vs_batched_max_abs": float(np.max(np.abs(g_sf - g_bat))),
           "measured_warm_batched_end_to_end": measured,
           "converged_record": {
               "couplings": len(rows),
               "runs_stopped_before_cap": {"plain": sum(r["stopped_plain"] for r in rows),
                                           "warm": sum(r["stopped_warm"] for r in rows)},
               "reached": {"plain": sum(r["iters_plain"] is not None for r in rows),
                           "warm": sum(r["iters_warm"] is not None for r in rows)},
               "median_iterations": {"plain": med(ip), "warm": med(iw)},
               "warm_start_own_gain_iterations_median": round(med(ip) / med(iw), 2),
               "median_time_to_dE_s": {"plain_sequential_ps_modelled": round(med(tp_model), 1),
                                       "warm_batched_modelled": round(med(tw_model), 2)},
               "speedup_median_modelled": round(med(tp_model) / med(tw_model), 1),
               "speedup_median_modelled_vs_plain_fused_sequential":
                   round(med([k * t_seqf for k in ip]) / med(tw_model), 1),
               "note": "plain times modelled: recorded iterations x measured per-iteration time; unreached targets "
                       "count as the cap"},
           "warm_start_own_gain_adjoint_measured": {
               "couplings": len(own), "runs": own,
               "median_s": {"plain": med([o["plain_s"] for o in own]), "warm": med([o["warm_s"] for o in own])},
               "speedup_median": round(med([o["plain_s"] for o in own]) / med([o["warm_s"] for o in own]), 2)}}
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
    if a.workload != "brickwork":
        # one side workload alone (for profiling it)
        run = {"uneven": lambda: uneven_batch_bench(lq),
               "gradient": lambda: ps_gradient_bench(lq, cpu=not a.no_cpu)}[a.workload]
        print(json.dumps({"workload": a.workload, "result": run()}))
        return
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
             # the circuit's two floors: its passes' reads and writes at the
             # measured copy rate, and its gates' FP64 work at the vector peak
             "hbm_floor_ms": round(passes * 2 * c * N / copy_gbs / 1e6, 2),
             "pass_kernel": "lq_pass (specialised)" if jit_launches else "k_tileg (generic)",
             "specialised_kernels": jit_kernels, "specialised_compile_s": round(jit_compile_s, 2),
             "specialised_launches": jit_launches}
    flops = algorithmic_flops(descs, a.qubits) if dtype == lq.C128 else None
    if flops is not None:
        fp64_floor_ms = flops / (FP64_PEAK_GFLOPS * 1e9) * 1e3
        fused["fp64_floor_ms"] = round(fp64_floor_ms, 2)
        # the same gates in VALU issue slots: each multiply takes a slot of
        # the FMA rate, so a rotation is 4 slots (8 flop-equivalents) per
        # amplitude
        fused["fp64_issue_floor_ms"] = round(fp64_floor_ms * 8 / 6, 2)
        fused["frac_of_floor"] = round(max(fused["hbm_floor_ms"], fp64_floor_ms) / (fused_s * 1e3), 4)
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
