"""Learned warm start for VQE: a small network from Hamiltonian couplings to
ansatz parameters, trained and evaluated on the GPU.

No counterpart in the reference: its VQE::run starts from the parameters the
caller passes (src/optimization/vqe.rs:55-60) and VQE::run_random from
uniform random ones (vqe.rs:102-111).  The network here only proposes a
starting point; `choose_start` keeps it only when the simulator's f64
energy at the proposal beats the plain initialisation.
"""
import ctypes
import json
import os

import numpy as np

from . import _lib


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class WarmStartNet:
    """Perceptron in -> hidden -> hidden -> out, GELU (erf form), bf16
    operands with f32 accumulation on the matrix cores, f32 master weights
    and Adam on the device (lq_mlp_* in the C ABI)."""

    def __init__(self, in_dim=3, hidden=256, out_dim=320, seed=0):
        self.in_dim, self.hidden, self.out_dim = int(in_dim), int(hidden), int(out_dim)
        h = ctypes.c_void_p()
        _lib.call("lq_mlp_create", self.in_dim, self.hidden, self.out_dim, int(seed), ctypes.byref(h))
        self._h = h
        n = ctypes.c_int64()
        _lib.call("lq_mlp_num_params", self._h, ctypes.byref(n))
        self.num_params = int(n.value)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.call("lq_mlp_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        out = np.empty(self.num_params, dtype=np.float32)
        _lib.call("lq_mlp_get_params", self._h, _fptr(out), self.num_params)
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _lib.call("lq_mlp_set_params", self._h, _fptr(p), int(p.size))

    def _x(self, x):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        if x.shape[1] != self.in_dim:
            raise ValueError("inputs must have %d columns" % self.in_dim)
        return x

    def forward(self, x):
        x = self._x(x)
        y = np.empty((x.shape[0], self.out_dim), dtype=np.float32)
        _lib.call("lq_mlp_forward", self._h, _fptr(x), x.shape[0], _fptr(y))
        return y

    def gradients(self, x, target):
        x = self._x(x)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(x.shape[0], self.out_dim)
        loss = np.zeros(1, dtype=np.float32)
        g = np.empty(self.num_params, dtype=np.float32)
        _lib.call("lq_mlp_gradients", self._h, _fptr(x), _fptr(t), x.shape[0], _fptr(loss), _fptr(g))
        return float(loss[0]), g

    def train(self, x, target, epochs, batch=256, learning_rate=1e-3, seed=0):
        x = self._x(x)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(x.shape[0], self.out_dim)
        hist = np.zeros(max(int(epochs), 1), dtype=np.float32)
        _lib.call("lq_mlp_train", self._h, _fptr(x), _fptr(t), x.shape[0], int(batch), int(epochs),
                  float(learning_rate), int(seed), _fptr(hist))
        return hist[:int(epochs)]


def choose_start(energy_fn, proposal, plain):
    """Return (start, used_proposal): the proposal when its energy (f64,
    from the simulator) is below the plain initialisation's, else plain."""
    ep = energy_fn(np.asarray(proposal, dtype=np.float64))
    e0 = energy_fn(np.asarray(plain, dtype=np.float64))
    if np.isfinite(ep) and ep < e0:
        return np.asarray(proposal, dtype=np.float64).copy(), True
    return np.asarray(plain, dtype=np.float64).copy(), False


# ---- the warm-start workload -------------------------------------------------------
# XYZ Heisenberg rings H(J) = sum_i Jx X_iX_{i+1} + Jy Y_iY_{i+1} + Jz Z_iZ_{i+1}
# (periodic), J ~ U[-1, 1]^3, with the RY/RZ + CZ-ring ansatz.  The plain
# initialisation is one fixed point theta0 ~ U[0, 2pi)^P (seeded), shared by
# every coupling; the network predicts theta*(J) - theta0 (scaled).

def couplings(count, seed):
    """count couplings (Jx, Jy, Jz) ~ U[-1, 1)^3 from SplitMix64(seed)."""
    from .algorithms import SplitMix64
    r = SplitMix64(seed)
    return np.array([[2.0 * r.uniform() - 1.0 for _ in range(3)] for _ in range(count)])


def plain_init(num_params, seed):
    """The plain initialisation theta0 ~ U[0, 2pi)^P from SplitMix64(seed)."""
    from .algorithms import SplitMix64
    r = SplitMix64(seed ^ 0x5A5A5A5A)
    return np.array([2.0 * np.pi * r.uniform() for _ in range(num_params)])


def hamiltonians(num_qubits, J):
    from .optimization import heisenberg_ring
    return [heisenberg_ring(num_qubits, *map(float, j)) for j in np.atleast_2d(J)]


def solve(ansatz, J, starts, optimizer, dtype=None, batch=64, log=None):
    """VQE::run (Adam, adjoint gradient) for each coupling J[i] from
    starts[i], in lock-step batches on the GPU.  Returns (params, energies,
    histories, converged)."""
    from .optimization import vqe_adam_batch
    from .states import C128
    dtype = C128 if dtype is None else dtype
    n = ansatz.num_qubits()
    J = np.atleast_2d(J)
    starts = np.atleast_2d(starts)
    xs, es, hs, cs = [], [], [], []
    for b0 in range(0, len(J), batch):
        res, conv = vqe_adam_batch(ansatz, hamiltonians(n, J[b0:b0 + batch]), starts[b0:b0 + batch], optimizer,
                                   dtype)
        for r in res:
            xs.append(np.array(r.optimal_parameters))
            es.append(r.ground_state_energy)
            hs.append(np.array(r.convergence_history))
        cs += list(conv)
        if log:
            log("solved %d of %d" % (len(xs), len(J)))
    return np.array(xs), np.array(es), hs, np.array(cs)


def generate_by_continuation(ansatz, J, theta0, optimizer, batch=64, seeds=None, log=None, state=None,
                             time_limit=None, checkpoint=None, optimizer_continue=None):
    """Converged VQE solutions for every coupling in J, grown outward from a
    few seeds: the seeds start from theta0; after that each round takes the
    `batch` unsolved couplings closest to the solved set and starts each from
    the solution of its nearest solved neighbour, so the solutions vary
    smoothly with J.  `state` (solved, X, E, iters) resumes an earlier run;
    `time_limit` (s) stops after the round that crosses it; `checkpoint` is
    called with the state after every round; `optimizer_continue` (default
    `optimizer`) runs the rounds that start from a solved neighbour.
    Returns the state."""
    import time
    t_start = time.perf_counter()
    J = np.asarray(J, dtype=np.float64)
    count = len(J)
    P = len(theta0)
    if state is None:
        state = (np.zeros(count, dtype=bool), np.zeros((count, P)), np.zeros(count), np.zeros(count, dtype=np.int64))
    solved, X, E, iters = state
    rnd = 0
    while not solved.all():
        if not solved.any():
            # seeds: farthest-point sampling from coupling 0, started from theta0
            k = max(1, min(batch if seeds is None else seeds, count))
            first = [0]
            d = np.linalg.norm(J - J[0], axis=1)
            while len(first) < k:
                j = int(np.argmax(d))
                first.append(j)
                d = np.minimum(d, np.linalg.norm(J - J[j], axis=1))
            take = np.array(first)
            st = np.tile(theta0, (len(take), 1))
        else:
            rest = np.flatnonzero(~solved)
            S = np.flatnonzero(solved)
            near = np.empty(len(rest))
            for c0 in range(0, len(rest), 1024):
                near[c0:c0 + 1024] = np.linalg.norm(J[rest[c0:c0 + 1024]][:, None, :] - J[S][None, :, :],
                                                    axis=2).min(axis=1)
            take = rest[np.argsort(near, kind="stable")[:batch]]
            nb = S[np.argmin(np.linalg.norm(J[take][:, None, :] - J[S][None, :, :], axis=2), axis=1)]
            st = X[nb]
        opt = optimizer if not solved.any() or optimizer_continue is None else optimizer_continue
        x, e, h, _ = solve(ansatz, J[take], st, opt, batch=batch)
        X[take], E[take] = x, e
        iters[take] = [len(t) for t in h]
        solved[take] = True
        if log:
            log("round %d: %d solved (%d total), mean iterations %.1f, %.1f s" % (
                rnd, len(take), solved.sum(), np.mean(iters[take]), time.perf_counter() - t_start))
        if checkpoint:
            checkpoint((solved, X, E, iters))
        rnd += 1
        if time_limit is not None and time.perf_counter() - t_start > time_limit:
            break
    return solved, X, E, iters


class WarmStart:
    """Trained warm start: J -> theta0 + scale * net(J)."""

    def __init__(self, net, theta0, scale):
        self.net, self.theta0, self.scale = net, np.asarray(theta0, dtype=np.float64), float(scale)

    # Adam learning rate per stage and its epochs (the recipe the 20-qubit
    # evaluation used: one stage)
    SCHEDULE = ((1e-3, 300),)

    @classmethod
    def train(cls, J, X, theta0, hidden=256, epochs=None, batch=256, learning_rate=None, seed=0, schedule=None):
        """Fit the network on the GPU to the converged parameters X (rows)
        of couplings J: Adam on the mean squared error, in stages of
        (learning rate, epochs) (`schedule`, default SCHEDULE; `epochs` and
        `learning_rate` together give one stage).  Returns (WarmStart, loss
        history per epoch)."""
        D = np.asarray(X, dtype=np.float64) - np.asarray(theta0)[None, :]
        scale = float(np.std(D)) or 1.0
        net = WarmStartNet(3, hidden, D.shape[1], seed=seed)
        if schedule is None:
            schedule = ((learning_rate, epochs),) if epochs is not None else cls.SCHEDULE
        xs, ts = np.asarray(J, dtype=np.float32), (D / scale).astype(np.float32)
        hist = []
        for k, (lr, ep) in enumerate(schedule):
            hist.append(net.train(xs, ts, int(ep), batch, float(lr if lr is not None else 1e-3), seed + k))
        return cls(net, theta0, scale), np.concatenate(hist) if hist else np.zeros(0)

    def predict(self, J):
        y = self.net.forward(np.asarray(np.atleast_2d(J), dtype=np.float32)).astype(np.float64)
        return self.theta0[None, :] + self.scale * y

    def get_state(self):
        return {"params": self.net.get_params(), "theta0": self.theta0, "scale": self.scale,
                "shape": np.array([self.net.in_dim, self.net.hidden, self.net.out_dim])}

    @classmethod
    def from_state(cls, st):
        i, h, o = (int(v) for v in st["shape"])
        net = WarmStartNet(i, h, o)
        net.set_params(np.asarray(st["params"], dtype=np.float32))
        return cls(net, st["theta0"], float(st["scale"]))


def iterations_to(history, target):
    """Index+1 of the first history entry <= target (None if never)."""
    h = np.asarray(history)
    k = np.flatnonzero(h <= target)
    return int(k[0]) + 1 if len(k) else None


def choose_starts(ws, ansatz, J, dtype=None):
    """Per coupling, the network's proposal when the simulator's f64
    energy there is below the plain initialisation's, else theta0.
    Returns (starts, used_proposal mask, proposal energies, plain energies)."""
    from .optimization import AdjointGradient
    from .states import C128
    dtype = C128 if dtype is None else dtype
    J = np.atleast_2d(J)
    prop = ws.predict(J)
    plain = np.tile(ws.theta0, (len(J), 1))
    obs = hamiltonians(ansatz.num_qubits(), J)
    adj = AdjointGradient(dtype)
    ep, _ = adj.compute_gradient_batch(ansatz, obs, prop, energy=True)
    e0, _ = adj.compute_gradient_batch(ansatz, obs, plain, energy=True)
    use = np.isfinite(ep) & (ep < e0)
    return np.where(use[:, None], prop, plain), use, ep, e0


# ---- command line: data generation and evaluation on the GPU --------------------------
WORKLOAD = {"num_qubits": 20, "layers": 8, "train": 4096, "held_out": 256, "seed": 5,
            "learning_rate": 0.05, "max_iterations": 600, "tolerance": 1e-8,
            # runs that start from a solved neighbour's parameters; the first
            # 993 training couplings were generated with 600 here
            "continue_max_iterations": 250}


def _save(path, state, J, theta0, meta):
    solved, X, E, iters = state
    tmp = path + ".tmp.npz"
    np.savez_compressed(tmp, J=J, X=X.astype(np.float32), E=E, iters=iters, solved=solved, theta0=theta0,
                        meta=np.array(json.dumps(meta)))
    os.replace(tmp, path)


def load_data(path):
    d = np.load(path, allow_pickle=False)
    meta = json.loads(str(d["meta"]))
    return d["J"], d["X"].astype(np.float64), d["E"], d["iters"], d["solved"], d["theta0"], meta


def cmd_generate(a):
    """Converged solutions for the training couplings (resumable)."""
    from .optimization import Adam, ry_rz_cz_ring_ansatz
    w = dict(WORKLOAD)
    ans = ry_rz_cz_ring_ansatz(w["num_qubits"], w["layers"])
    J = couplings(w["train"], w["seed"])
    theta0 = plain_init(ans.num_parameters(), w["seed"])
    state = None
    if os.path.exists(a.out):
        Jd, X, E, it, solved, th, meta = load_data(a.out)
        if not (np.array_equal(Jd, J) and np.array_equal(th, theta0)):
            raise SystemExit("existing %s holds a different workload" % a.out)
        state = (solved, X, E, it)
        print("resuming: %d of %d solved" % (solved.sum(), len(J)), flush=True)
    opt = Adam(w["learning_rate"], w["max_iterations"]).with_tolerance(w["tolerance"])
    opt_c = Adam(w["learning_rate"], w["continue_max_iterations"]).with_tolerance(w["tolerance"])
    state = generate_by_continuation(ans, J, theta0, opt, batch=a.batch, seeds=1, optimizer_continue=opt_c,
                                     log=lambda m: print(m, flush=True),
                                     state=state, time_limit=a.time_limit,
                                     checkpoint=lambda s: _save(a.out, s, J, theta0, w))
    print("done: %d of %d solved" % (state[0].sum(), len(J)), flush=True)


def time_to_target(histories, targets, cap):
    """Iterations until each history is within its target (cap if never)."""
    out = []
    for h, t in zip(histories, targets):
        k = iterations_to(h, t)
        out.append(cap if k is None else k)
    return np.array(out)


def pick_learning_rate(ans, J, starts, grid, max_iterations, tolerance, dtype=None, batch=64):
    """Adam learning rate (from `grid`) with the lowest median iterations to
    within 1e-4 of the best energy any run reached, per coupling.  Returns
    (best rate, {rate: median iterations})."""
    from .optimization import Adam
    runs = {}
    for lr in grid:
        _, e, h, _ = solve(ans, J, starts, Adam(lr, max_iterations).with_tolerance(tolerance), dtype, batch)
        runs[lr] = (e, h)
    e_best = np.min(np.array([runs[lr][0] for lr in grid]), axis=0)
    med = {lr: float(np.median(time_to_target(runs[lr][1], e_best + 1e-4, max_iterations))) for lr in grid}
    return min(grid, key=lambda lr: (med[lr], lr)), med


def evaluate(ans, ws, J, theta0, lr_plain, lr_warm, max_iterations, tolerance, batch=64, dtype=None, log=None):
    """Plain and warm-started VQE::run (Adam, adjoint gradient: the same
    trajectory as parameter shift) on couplings J.  Per coupling: iterations
    until the energy is within 1e-4 of the lower of the two converged
    energies (None: not reached), and the start energies."""
    from .optimization import Adam
    starts, used, ep, e0 = choose_starts(ws, ans, J, dtype)
    _, e_plain, h_plain, _ = solve(ans, J, np.tile(theta0, (len(J), 1)),
                                   Adam(lr_plain, max_iterations).with_tolerance(tolerance), dtype, batch,
                                   log and (lambda m: log("plain: " + m)))
    _, e_warm, h_warm, _ = solve(ans, J, starts, Adam(lr_warm, max_iterations).with_tolerance(tolerance), dtype,
                                 batch, log and (lambda m: log("warm: " + m)))
    rows = []
    for i in range(len(J)):
        e_star = min(e_plain[i], e_warm[i])
        rows.append({"J": [float(v) for v in J[i]], "e_star": float(e_star),
                     "e_plain_final": float(e_plain[i]), "e_warm_final": float(e_warm[i]),
                     "e_start_plain": float(e0[i]), "e_start_proposal": float(ep[i]), "used_proposal": bool(used[i]),
                     "iters_plain": iterations_to(h_plain[i], e_star + 1e-4),
                     "iters_warm": iterations_to(h_warm[i], e_star + 1e-4),
                     # against the plain method's own converged energy
                     "iters_plain_to_plain": iterations_to(h_plain[i], e_plain[i] + 1e-4),
                     "iters_warm_to_plain": iterations_to(h_warm[i], e_plain[i] + 1e-4),
                     "len_plain": len(h_plain[i]), "len_warm": len(h_warm[i])})
    return rows


def cmd_evaluate(a):
    from .optimization import Adam, ry_rz_cz_ring_ansatz
    J, X, E, it, solved, theta0, meta = load_data(a.data)
    if not solved.all():
        raise SystemExit("training data incomplete: %d of %d solved" % (solved.sum(), len(J)))
    ans = ry_rz_cz_ring_ansatz(meta["num_qubits"], meta["layers"])
    ws, hist = WarmStart.train(J, X, theta0, epochs=a.epochs, learning_rate=a.lr)
    log0 = lambda m: print(m, flush=True)
    log0("trained %d epochs" % len(hist))
    print("trained: loss %.4g -> %.4g" % (hist[0], hist[-1]), flush=True)
    Jte = couplings(meta["train"] + meta["held_out"], meta["seed"])[meta["train"]:][a.offset:a.offset + a.count]
    log = lambda m: print(m, flush=True)
    if a.lr_plain and a.lr_warm:
        lr_p, lr_w, med_p, med_w = a.lr_plain, a.lr_warm, None, None
    else:
        # learning rate per method on validation couplings (neither training nor held out)
        Jval = couplings(a.validation, meta["seed"] + 1)
        grid = tuple(float(v) for v in a.lr_grid.split(","))
        lr_p, med_p = pick_learning_rate(ans, Jval, np.tile(theta0, (len(Jval), 1)), grid, a.max_iterations,
                                         a.tolerance, batch=a.batch)
        log("plain learning rate %g %s" % (lr_p, med_p))
        lr_w, med_w = pick_learning_rate(ans, Jval, choose_starts(ws, ans, Jval)[0], grid, a.max_iterations,
                                         a.tolerance, batch=a.batch)
        log("warm learning rate %g %s" % (lr_w, med_w))
    rows = evaluate(ans, ws, Jte, theta0, lr_p, lr_w, a.max_iterations, a.tolerance, batch=a.batch, log=log)
    with open(a.out, "w") as f:
        json.dump({"meta": meta, "train_loss": [float(v) for v in hist], "lr_plain": lr_p, "lr_warm": lr_w,
                   "epochs": a.epochs, "lr": a.lr, "max_iterations": a.max_iterations, "tolerance": a.tolerance,
                   "validation_median_iterations": {"plain": med_p, "warm": med_w}, "offset": a.offset,
                   "rows": rows}, f)
    ip = [r["iters_plain"] for r in rows]
    iw = [r["iters_warm"] for r in rows]
    print("held-out %d: plain reached %d, warm reached %d; median iterations plain %s warm %s" % (
        len(rows), sum(v is not None for v in ip), sum(v is not None for v in iw),
        np.median([v for v in ip if v is not None]) if any(v is not None for v in ip) else None,
        np.median([v for v in iw if v is not None]) if any(v is not None for v in iw) else None), flush=True)


def cmd_converge(a):
    """E* by runs to convergence: for the first `count` held-out couplings,
    VQE::run with Adam (adjoint gradient) from the plain and the warm starts
    at the evaluation's learning rates, up to `max_iterations` iterations at
    `tolerance`; E* is the lower of the two final energies.  Records per
    coupling E*, both final energies and iteration counts to E* + 1e-4, and
    whether each run stopped by Adam's own tests before the cap."""
    from .optimization import Adam, ry_rz_cz_ring_ansatz
    J, X, E, it, solved, theta0, meta = load_data(a.data)
    with open(a.eval) as f:
        ev = json.load(f)
    ans = ry_rz_cz_ring_ansatz(meta["num_qubits"], meta["layers"])
    ws, hist = WarmStart.train(J, X, theta0, epochs=ev.get("epochs"), learning_rate=ev.get("lr"))
    print("trained: loss %.4g -> %.4g" % (hist[0], hist[-1]), flush=True)
    Jte = couplings(meta["train"] + meta["held_out"], meta["seed"])[meta["train"]:][:a.count]
    starts, used, ep, e0 = choose_starts(ws, ans, Jte)
    lr_p, lr_w = ev["lr_plain"], ev["lr_warm"]
    log = lambda m: print(m, flush=True)
    _, e_p, h_p, c_p = solve(ans, Jte, np.tile(theta0, (len(Jte), 1)),
                             Adam(lr_p, a.max_iterations).with_tolerance(a.tolerance), batch=a.batch,
                             log=lambda m: log("plain: " + m))
    _, e_w, h_w, c_w = solve(ans, Jte, starts, Adam(lr_w, a.max_iterations).with_tolerance(a.tolerance),
                             batch=a.batch, log=lambda m: log("warm: " + m))
    rows = []
    for i in range(len(Jte)):
        e_star = float(min(e_p[i], e_w[i]))
        rows.append({"J": [float(v) for v in Jte[i]], "e_star": e_star, "e_plain_final": float(e_p[i]),
                     "e_warm_final": float(e_w[i]), "used_proposal": bool(used[i]),
                     "iters_plain": iterations_to(h_p[i], e_star + 1e-4),
                     "iters_warm": iterations_to(h_w[i], e_star + 1e-4),
                     "len_plain": len(h_p[i]), "len_warm": len(h_w[i]),
                     "stopped_plain": bool(c_p[i]), "stopped_warm": bool(c_w[i])})
    with open(a.out, "w") as f:
        json.dump({"meta": meta, "lr_plain": lr_p, "lr_warm": lr_w, "epochs": ev.get("epochs"), "lr": ev.get("lr"),
                   "max_iterations": a.max_iterations, "tolerance": a.tolerance, "count": len(rows), "rows": rows}, f)
    ip = [r["iters_plain"] for r in rows]
    iw = [r["iters_warm"] for r in rows]
    print("converged %d: plain reached %d (median %s), warm reached %d (median %s); stopped before the cap: "
          "plain %d, warm %d" % (len(rows), sum(v is not None for v in ip),
                                 np.median([v or a.max_iterations for v in ip]), sum(v is not None for v in iw),
                                 np.median([v or a.max_iterations for v in iw]), sum(r["stopped_plain"] for r in rows),
                                 sum(r["stopped_warm"] for r in rows)), flush=True)


def cmd_merge(a):
    """Join evaluation parts (disjoint --offset ranges of the held-out
    couplings, same training recipe and learning rates) into one file."""
    parts = []
    for path in a.parts:
        with open(path) as f:
            parts.append(json.load(f))
    keys = ("lr_plain", "lr_warm", "epochs", "lr", "max_iterations", "tolerance")
    for p in parts[1:]:
        for k in keys:
            if p[k] != parts[0][k]:
                raise SystemExit("parts differ in %s: %r vs %r" % (k, p[k], parts[0][k]))
    parts.sort(key=lambda p: p["offset"])
    rows, end = [], 0
    for p in parts:
        if p["offset"] != end:
            raise SystemExit("parts do not tile the held-out couplings: offset %d after %d" % (p["offset"], end))
        rows += p["rows"]
        end += len(p["rows"])
    out = dict(parts[0])
    out["rows"] = rows
    out["parts"] = [{"offset": p["offset"], "count": len(p["rows"]),
                     "validation_median_iterations": p["validation_median_iterations"]} for p in parts]
    with open(a.out, "w") as f:
        json.dump(out, f)
    print("merged %d parts, %d couplings -> %s" % (len(parts), len(rows), a.out))


def cmd_time_gradient(a):
    """Wall time of single gradients of the benchmark ansatz (one XYZ ring
    coupling, the plain initialisation), one line per call."""
    import time
    from .optimization import ParameterShift, AdjointGradient, ry_rz_cz_ring_ansatz, heisenberg_ring
    from ._lib import call
    if a.tile_bits:
        call("lq_set_fused_config", a.tile_bits, a.tile_bits + 1, a.low_bits)
    call("lq_set_expect_tiles", {"off": 0, "auto": 1, "on": 2}[a.expect_tiles])
    call("lq_set_jit", {"off": 0, "repeat": 1, "always": 2}[a.jit])
    ans = ry_rz_cz_ring_ansatz(a.qubits, a.layers)
    obs = heisenberg_ring(a.qubits, 0.3, -0.7, 0.5)
    theta = plain_init(ans.num_parameters(), 5)
    method = {"batched": ParameterShift(batched=True), "sequential": ParameterShift(batched=False),
              "adjoint": AdjointGradient()}[a.method]
    for i in range(a.calls):
        t = time.perf_counter()
        e, _ = method.compute_gradient(ans, obs, theta, energy=True)
        call("lq_synchronize")
        print("call %d: %.4f s (energy %.12f)" % (i, time.perf_counter() - t, e), flush=True)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(prog="python -m logosq_amd.warmstart")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate", help="converged VQE solutions of the training couplings")
    g.add_argument("--out", required=True)
    g.add_argument("--batch", type=int, default=16)
    g.add_argument("--time-limit", type=float, default=None)
    e = sub.add_parser("evaluate", help="train on the GPU, then warm vs plain on held-out couplings")
    e.add_argument("--data", required=True)
    e.add_argument("--out", required=True)
    e.add_argument("--count", type=int, default=256)
    e.add_argument("--epochs", type=int, default=None, help="one stage of this many epochs (default: the schedule)")
    e.add_argument("--lr", type=float, default=None)
    e.add_argument("--max-iterations", type=int, default=1500)
    e.add_argument("--tolerance", type=float, default=1e-10)
    e.add_argument("--validation", type=int, default=16)
    e.add_argument("--offset", type=int, default=0)
    e.add_argument("--batch", type=int, default=16)
    e.add_argument("--lr-plain", type=float, default=None)
    e.add_argument("--lr-warm", type=float, default=None)
    e.add_argument("--lr-grid", default="0.02,0.05,0.1")
    c = sub.add_parser("converge", help="E* of held-out couplings from runs to convergence, both starts")
    c.add_argument("--data", required=True)
    c.add_argument("--eval", required=True, help="the evaluation record (learning rates, training recipe)")
    c.add_argument("--out", required=True)
    c.add_argument("--count", type=int, default=32)
    c.add_argument("--max-iterations", type=int, default=5000)
    c.add_argument("--tolerance", type=float, default=1e-12)
    c.add_argument("--batch", type=int, default=16)
    m = sub.add_parser("merge", help="join evaluation parts into one file")
    m.add_argument("--out", required=True)
    m.add_argument("parts", nargs="+")
    t = sub.add_parser("time-gradient", help="wall time of gradients of the benchmark ansatz")
    t.add_argument("--qubits", type=int, default=20)
    t.add_argument("--layers", type=int, default=8)
    t.add_argument("--method", choices=["batched", "sequential", "adjoint"], default="batched")
    t.add_argument("--calls", type=int, default=3)
    t.add_argument("--tile-bits", type=int, default=0, help="fused tile of 2^k complex128 amplitudes (default: 12)")
    t.add_argument("--low-bits", type=int, default=5)
    t.add_argument("--expect-tiles", choices=["off", "auto", "on"], default="auto",
                   help="flip groups of the expectation from LDS tiles (lq_set_expect_tiles)")
    t.add_argument("--jit", choices=["off", "repeat", "always"], default="repeat",
                   help="run-time specialised tile passes (lq_set_jit)")
    a = ap.parse_args(argv)
    {"generate": cmd_generate, "evaluate": cmd_evaluate, "converge": cmd_converge, "merge": cmd_merge,
     "time-gradient": cmd_time_gradient}[a.cmd](a)


if __name__ == "__main__":
    main()
