"""Optimization: mirrors of src/optimization/{observable,ansatz,gradient,
optimizer,vqe}.rs on top of the GPU path.

Observables compile to Pauli-code matrices for lq_expect_pauli_sum; ansatzes
expose a gate *template* (descriptors with parameter slots) so the gradient
and VQE loops run inside the native library.
"""
import ctypes
import enum
import math

import numpy as np

from . import abi
from ._lib import call, load
from .circuits import Circuit, Param
from .states import State, C128


# ---- observables (observable.rs) ---------------------------------------------
class Pauli(enum.IntEnum):
    I = abi.PAULI_I
    X = abi.PAULI_X
    Y = abi.PAULI_Y
    Z = abi.PAULI_Z

    def matrix(self):
        return {0: np.eye(2, dtype=complex), 1: np.array([[0, 1], [1, 0]], dtype=complex),
                2: np.array([[0, -1j], [1j, 0]]), 3: np.array([[1, 0], [0, -1]], dtype=complex)}[int(self)]


class Observable:
    """trait Observable (observable.rs:10-18)."""

    def expectation(self, state):
        raise NotImplementedError

    def matrix(self):
        return None


class PauliTerm:
    """PauliTerm (observable.rs:79-148)."""

    def __init__(self, coefficient, paulis):
        self.coefficient = float(coefficient)
        self.paulis = [Pauli(int(p)) for p in paulis]

    def expectation(self, state, convention=abi.PAULI_REFERENCE):
        n = state.num_qubits()
        if len(self.paulis) != n:
            raise AssertionError("Pauli term must have operator for each qubit")
        return _expect(state, np.array([self.paulis], dtype=np.uint8), np.array([self.coefficient]), convention)


def _expect(state, codes, coeffs, convention):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
    out = ctypes.c_double()
    call("lq_expect_pauli_sum", state.handle, codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
         coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(coeffs), convention, ctypes.byref(out))
    return out.value


class PauliObservable(Observable):
    """PauliObservable (observable.rs:152-232).  `convention` selects the
    reference's row-bit Y phase (default) or the physical one."""

    def __init__(self, num_qubits, convention=abi.PAULI_REFERENCE):
        self.terms = []
        self.num_qubits = int(num_qubits)
        self.convention = convention

    def add_term(self, term):
        if len(term.paulis) != self.num_qubits:
            raise AssertionError("Term must match observable qubit count")
        self.terms.append(term)

    @classmethod
    def single_z(cls, num_qubits, target_qubit):
        if not target_qubit < num_qubits:
            raise AssertionError("Qubit index out of bounds")
        p = [Pauli.I] * num_qubits
        p[target_qubit] = Pauli.Z
        o = cls(num_qubits)
        o.add_term(PauliTerm(1.0, p))
        return o

    @classmethod
    def zz_interaction(cls, num_qubits, qubit1, qubit2, coefficient):
        if not qubit1 < num_qubits or not qubit2 < num_qubits:
            raise AssertionError("Qubit index out of bounds")
        p = [Pauli.I] * num_qubits
        p[qubit1] = Pauli.Z
        p[qubit2] = Pauli.Z
        o = cls(num_qubits)
        o.add_term(PauliTerm(coefficient, p))
        return o

    @classmethod
    def ising_hamiltonian(cls, num_qubits, j_couplings, h_fields):
        o = cls(num_qubits)
        for i, j in enumerate(j_couplings):
            if not i + 1 < num_qubits:
                raise AssertionError("Qubit index out of bounds")
            p = [Pauli.I] * num_qubits
            p[i] = Pauli.Z
            p[i + 1] = Pauli.Z
            o.add_term(PauliTerm(j, p))
        for i, hv in enumerate(h_fields):
            p = [Pauli.I] * num_qubits
            p[i] = Pauli.Z
            o.add_term(PauliTerm(hv, p))
        return o

    def codes(self):
        if not self.terms:
            return np.zeros((0, self.num_qubits), dtype=np.uint8), np.zeros(0)
        return (np.array([[int(p) for p in t.paulis] for t in self.terms], dtype=np.uint8),
                np.array([t.coefficient for t in self.terms], dtype=np.float64))

    def expectation(self, state):
        codes, coeffs = self.codes()
        if len(coeffs) == 0:
            return 0.0
        return _expect(state, codes, coeffs, self.convention)


def heisenberg_ring(num_qubits, jx, jy, jz, convention=abi.PAULI_REFERENCE):
    """XYZ Heisenberg ring H = sum_i Jx X_iX_{i+1} + Jy Y_iY_{i+1} + Jz Z_iZ_{i+1}
    (periodic), as a PauliObservable with 3n terms ordered bond by bond."""
    o = PauliObservable(num_qubits, convention)
    for i in range(num_qubits):
        j = (i + 1) % num_qubits
        for pa, c in ((Pauli.X, jx), (Pauli.Y, jy), (Pauli.Z, jz)):
            p = [Pauli.I] * num_qubits
            p[i] = pa
            p[j] = pa
            o.add_term(PauliTerm(c, p))
    return o


# ---- ansatzes (ansatz.rs) --------------------------------------------------------
class Ansatz:
    """trait Ansatz (ansatz.rs:10-43) with a template for the native loops."""

    def build_circuit(self, parameters):
        raise NotImplementedError

    def num_parameters(self):
        raise NotImplementedError

    def num_qubits(self):
        raise NotImplementedError

    def apply(self, state, parameters):
        if state.num_qubits() != self.num_qubits():
            raise AssertionError("State qubit count must match ansatz")
        if len(parameters) != self.num_parameters():
            raise AssertionError("Parameter count mismatch")
        self.build_circuit(parameters).execute(state)

    def parameter_bounds(self):
        return None

    def template(self):
        """Gate descriptors with parameter slots: build the circuit on Param refs."""
        c = self.build_circuit([Param(i, 0.0) for i in range(self.num_parameters())])
        descs = c.descriptors()
        if any(d is None for d in descs):
            return None
        return descs


class EntanglingGate:
    CNOT = ("CNOT", None)
    CZ = ("CZ", None)

    @staticmethod
    def RZZ(theta):
        return ("RZZ", float(theta))


class EntanglingPattern(enum.Enum):
    Linear = 0
    Circular = 1
    Full = 2


class HardwareEfficientAnsatz(Ansatz):
    """HardwareEfficientAnsatz (ansatz.rs:63-186)."""

    def __init__(self, num_qubits, depth, entangling_gate=EntanglingGate.CNOT,
                 entangling_pattern=EntanglingPattern.Linear):
        self.nq, self.depth = num_qubits, depth
        self.entangling_gate, self.entangling_pattern = entangling_gate, entangling_pattern

    def num_parameters(self):
        return self.nq * (self.depth + 1)

    def num_qubits(self):
        return self.nq

    def parameter_bounds(self):
        return [(0.0, 2.0 * math.pi)] * self.num_parameters()

    def _ent(self, c, i, j):
        kind, th = self.entangling_gate
        if kind == "CNOT":
            c.cnot(i, j)
        elif kind == "CZ":
            c.cz(i, j)
        else:
            c.rzz(i, j, th)

    def build_circuit(self, parameters):
        if len(parameters) != self.num_parameters():
            raise AssertionError("Parameter count mismatch")
        c = Circuit(self.nq)
        k = 0
        for layer in range(self.depth):
            for q in range(self.nq):
                c.ry(q, parameters[k])
                k += 1
            if layer < self.depth - 1 or self.depth == 1:
                if self.entangling_pattern == EntanglingPattern.Linear:
                    for i in range(self.nq - 1):
                        self._ent(c, i, i + 1)
                elif self.entangling_pattern == EntanglingPattern.Circular:
                    for i in range(self.nq - 1):
                        self._ent(c, i, i + 1)
                    if self.nq > 2:
                        self._ent(c, self.nq - 1, 0)
                else:
                    for i in range(self.nq):
                        for j in range(i + 1, self.nq):
                            self._ent(c, i, j)
        for q in range(self.nq):
            c.ry(q, parameters[k])
            k += 1
        return c


class RealAmplitudesAnsatz(Ansatz):
    """RealAmplitudesAnsatz (ansatz.rs:190-248)."""

    def __init__(self, num_qubits, depth):
        self.nq, self.depth = num_qubits, depth

    def num_parameters(self):
        if self.nq > 1 and self.depth == 1:
            return self.nq * (self.depth + 2)
        return self.nq * (self.depth + 1)

    def num_qubits(self):
        return self.nq

    def build_circuit(self, parameters):
        c = Circuit(self.nq)
        k = 0
        if self.nq > 1 and self.depth == 1:
            for q in range(self.nq):
                c.ry(q, parameters[k])
                k += 1
        for _ in range(self.depth):
            for q in range(self.nq):
                c.ry(q, parameters[k])
                k += 1
            for i in range(self.nq - 1):
                c.cnot(i, i + 1)
        for q in range(self.nq):
            c.ry(q, parameters[k])
            k += 1
        return c


class EfficientSU2Ansatz(Ansatz):
    """EfficientSU2Ansatz (ansatz.rs:252-295)."""

    def __init__(self, num_qubits, depth):
        self.nq, self.depth = num_qubits, depth

    def num_parameters(self):
        return 3 * self.nq * self.depth

    def num_qubits(self):
        return self.nq

    def build_circuit(self, parameters):
        c = Circuit(self.nq)
        k = 0
        for _ in range(self.depth):
            for q in range(self.nq):
                c.rz(q, parameters[k])
                c.ry(q, parameters[k + 1])
                c.rz(q, parameters[k + 2])
                k += 3
            for i in range(self.nq - 1):
                c.cnot(i, i + 1)
        return c


class ParameterizedCircuit(Ansatz):
    """ParameterizedCircuit (ansatz.rs:298-335): a builder closure."""

    def __init__(self, num_qubits, num_parameters, builder):
        self.nq, self.np_, self.builder = num_qubits, num_parameters, builder

    def num_parameters(self):
        return self.np_

    def num_qubits(self):
        return self.nq

    def build_circuit(self, parameters):
        return self.builder(parameters)

    def template(self):
        try:
            return super().template()
        except (TypeError, ValueError):
            return None


def ry_rz_cz_ring_ansatz(num_qubits, layers):
    """The benchmark ansatz: per layer RY(q) then RZ(q) on every qubit, then CZ
    on the ring (q, q+1 mod n); parameters ordered layer -> qubit -> (RY, RZ)."""

    def build(p):
        c = Circuit(num_qubits)
        k = 0
        for _ in range(layers):
            for q in range(num_qubits):
                c.ry(q, p[k])
                c.rz(q, p[k + 1])
                k += 2
            for q in range(num_qubits):
                c.cz(q, (q + 1) % num_qubits)
        return c

    return ParameterizedCircuit(num_qubits, 2 * num_qubits * layers, build)


# ---- gradients (gradient.rs) ------------------------------------------------------
def _template_args(ansatz):
    t = ansatz.template()
    if t is None:
        raise TypeError("ansatz has no gate template (dense MatrixGate ops or parameter arithmetic)")
    return abi.gate_array(t), len(t)


def _obs_args(observable):
    codes, coeffs = observable.codes()
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
    return codes, coeffs


class ParameterShift:
    """ParameterShift (gradient.rs:22-78).  batched=True runs all 2P+1 circuits
    together (tile passes) instead of one after another."""

    def __init__(self, shift=math.pi / 2.0, batched=False, dtype=C128):
        self.shift = shift
        self.batched = batched
        self.dtype = dtype

    @classmethod
    def new(cls):
        return cls()

    @classmethod
    def with_shift(cls, shift):
        return cls(shift)

    def compute_gradient(self, ansatz, observable, parameters, energy=False):
        arr, ng = _template_args(ansatz)
        codes, coeffs = _obs_args(observable)
        p = np.ascontiguousarray(parameters, dtype=np.float64)
        grad = np.zeros(len(p))
        e = ctypes.c_double()
        mode = abi.GRAD_BATCHED if self.batched else abi.GRAD_SEQUENTIAL
        call("lq_param_shift_gradient", ansatz.num_qubits(), self.dtype, arr, ng,
             p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(p), self.shift,
             codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
             coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(coeffs), observable.convention, mode,
             ctypes.byref(e) if energy else None, grad.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if energy:
            return e.value, grad.tolist()
        return grad.tolist()


def _energy(ansatz, observable, params):
    st = State.zero_state(ansatz.num_qubits())
    ansatz.apply(st, params)
    return observable.expectation(st)


class AdjointGradient:
    """Adjoint-method gradient on the GPU (no reference counterpart): equal to
    ParameterShift (shift pi/2) up to rounding for RX/RY/RZ parameters, at
    the cost of one forward and one backward sweep."""

    shift = math.pi / 2.0
    batched = False
    mode = abi.GRAD_ADJOINT

    def __init__(self, dtype=C128):
        self.dtype = dtype

    def compute_gradient(self, ansatz, observable, parameters, energy=False):
        arr, ng = _template_args(ansatz)
        codes, coeffs = _obs_args(observable)
        p = np.ascontiguousarray(parameters, dtype=np.float64)
        grad = np.zeros(len(p))
        e = ctypes.c_double()
        call("lq_adjoint_gradient", ansatz.num_qubits(), self.dtype, arr, ng,
             p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(p),
             codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
             coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(coeffs), observable.convention,
             ctypes.byref(e) if energy else None, grad.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if energy:
            return e.value, grad.tolist()
        return grad.tolist()


    def compute_gradient_batch(self, ansatz, observables, parameter_sets, energy=False):
        """Adjoint gradients of many (observable, parameters) pairs of one
        ansatz in one batched call; the observables must share their Pauli
        strings (only the coefficients differ)."""
        arr, ng = _template_args(ansatz)
        codes, coeff_sets = _shared_strings(observables)
        x = np.ascontiguousarray(parameter_sets, dtype=np.float64).reshape(len(observables), -1)
        grads = np.zeros_like(x)
        e = np.zeros(len(observables))
        pd = ctypes.POINTER(ctypes.c_double)
        call("lq_adjoint_gradient_batch", ansatz.num_qubits(), self.dtype, arr, ng, len(observables),
             x.ctypes.data_as(pd), x.shape[1], codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
             coeff_sets.ctypes.data_as(pd), coeff_sets.shape[1], observables[0].convention,
             e.ctypes.data_as(pd), grads.ctypes.data_as(pd))
        return (e, grads) if energy else grads


def _shared_strings(observables):
    """(codes, coefficient matrix) of observables with identical Pauli strings."""
    codes0, _ = _obs_args(observables[0])
    coeff_sets = []
    for o in observables:
        c, k = _obs_args(o)
        if c.shape != codes0.shape or not np.array_equal(c, codes0) or o.convention != observables[0].convention:
            raise ValueError("batched observables must share their Pauli strings and convention")
        coeff_sets.append(k)
    return codes0, np.ascontiguousarray(np.array(coeff_sets, dtype=np.float64).reshape(len(observables), -1))


def vqe_adam_batch(ansatz, hamiltonians, initial_parameters, optimizer, dtype=C128):
    """VQE::run with Adam and the adjoint gradient for many Hamiltonians of
    one family (shared Pauli strings) in lock step on the GPU; every run
    follows Adam::minimize exactly.  Returns a list of VQEResult (without
    final states) plus the converged flags."""
    arr, ng = _template_args(ansatz)
    codes, coeff_sets = _shared_strings(hamiltonians)
    count = len(hamiltonians)
    x = np.array(initial_parameters, dtype=np.float64, copy=True).reshape(count, -1)
    it_max = int(optimizer.max_iterations)
    hist = np.zeros((count, max(1, it_max)))
    e = np.zeros(count)
    iters = np.zeros(count, dtype=np.int32)
    conv = np.zeros(count, dtype=np.int32)
    pd = ctypes.POINTER(ctypes.c_double)
    pi = ctypes.POINTER(ctypes.c_int)
    call("lq_vqe_adam_batch", ansatz.num_qubits(), dtype, arr, ng, x.shape[1],
         codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), coeff_sets.ctypes.data_as(pd), coeff_sets.shape[1],
         hamiltonians[0].convention, count, optimizer.learning_rate, it_max, optimizer.tolerance,
         x.ctypes.data_as(pd), hist.ctypes.data_as(pd), e.ctypes.data_as(pd), iters.ctypes.data_as(pi),
         conv.ctypes.data_as(pi))
    out = [VQEResult(e[i], x[i].tolist(), hist[i, :iters[i]].tolist(), None) for i in range(count)]
    return out, conv.astype(bool)


class FiniteDifference:
    """FiniteDifference (gradient.rs:81-125): forward differences."""

    def __init__(self, epsilon=1e-7):
        self.epsilon = epsilon

    def compute_gradient(self, ansatz, observable, parameters):
        fx = _energy(ansatz, observable, parameters)
        g = []
        for i in range(len(parameters)):
            p = list(parameters)
            p[i] += self.epsilon
            g.append((_energy(ansatz, observable, p) - fx) / self.epsilon)
        return g


# ---- optimizers (optimizer.rs) and VQE (vqe.rs) ----------------------------------
class LqOptResult(ctypes.Structure):
    _fields_ = [("optimal_value", ctypes.c_double), ("num_iterations", ctypes.c_int32),
                ("converged", ctypes.c_int32)]


from ._lib import OBJ_FN, GRAD_FN  # noqa: E402


class OptimizerResult:
    def __init__(self, params, value, history, converged):
        self.optimal_parameters = params
        self.optimal_value = value
        self.num_iterations = len(history)
        self.convergence_history = history
        self.converged = converged


def _wrap(f, grad_f):
    def fo(x, dim, _u):
        return float(f([x[i] for i in range(dim)]))

    def go(x, dim, out, _u):
        g = grad_f([x[i] for i in range(dim)])
        for i in range(dim):
            out[i] = float(g[i])

    return OBJ_FN(fo), GRAD_FN(go)


class GradientDescent:
    def __init__(self, learning_rate, max_iterations, tolerance=1e-6):
        self.learning_rate, self.max_iterations, self.tolerance = learning_rate, max_iterations, tolerance

    def with_tolerance(self, tol):
        self.tolerance = tol
        return self

    def minimize(self, f, grad_f, initial_params):
        x = np.array(initial_params, dtype=np.float64, copy=True)  # the caller's slice is not modified
        hist = np.zeros(max(1, self.max_iterations))
        res = LqOptResult()
        fo, go = _wrap(f, grad_f)
        call("lq_gd_minimize", self.learning_rate, self.max_iterations, self.tolerance, fo, go, None,
             x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(x),
             hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(res))
        return OptimizerResult(x.tolist(), res.optimal_value, hist[:res.num_iterations].tolist(), bool(res.converged))


class Adam:
    def __init__(self, learning_rate, max_iterations, tolerance=1e-6):
        self.learning_rate, self.max_iterations, self.tolerance = learning_rate, max_iterations, tolerance
        self.beta1, self.beta2, self.epsilon = 0.9, 0.999, 1e-8

    def with_tolerance(self, tol):
        self.tolerance = tol
        return self

    def minimize(self, f, grad_f, initial_params):
        x = np.array(initial_params, dtype=np.float64, copy=True)  # the caller's slice is not modified
        hist = np.zeros(max(1, self.max_iterations))
        res = LqOptResult()
        fo, go = _wrap(f, grad_f)
        call("lq_adam_minimize", self.learning_rate, self.beta1, self.beta2, self.epsilon, self.max_iterations,
             self.tolerance, fo, go, None, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(x),
             hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(res))
        return OptimizerResult(x.tolist(), res.optimal_value, hist[:res.num_iterations].tolist(), bool(res.converged))


class VQEResult:
    def __init__(self, energy, params, history, final_state):
        self.ground_state_energy = energy
        self.optimal_parameters = params
        self.num_iterations = len(history)
        self.convergence_history = history
        self.final_state = final_state


class VQE:
    """VQE (vqe.rs:23-112): the whole loop runs in the native library."""

    def __init__(self, ansatz, hamiltonian, gradient_method, optimizer):
        self.ansatz, self.hamiltonian = ansatz, hamiltonian
        self.gradient_method, self.optimizer = gradient_method, optimizer
        self.verbose = True

    def run(self, initial_parameters):
        arr, ng = _template_args(self.ansatz)
        codes, coeffs = _obs_args(self.hamiltonian)
        x = np.array(initial_parameters, dtype=np.float64, copy=True)
        opt = self.optimizer
        hist = np.zeros(max(1, opt.max_iterations))
        res = LqOptResult()
        gm = self.gradient_method
        mode = getattr(gm, "mode", abi.GRAD_BATCHED if getattr(gm, "batched", False) else abi.GRAD_SEQUENTIAL)
        dtype = getattr(gm, "dtype", C128)
        final = State.zero_state(self.ansatz.num_qubits(), dtype)
        call("lq_vqe_run", self.ansatz.num_qubits(), dtype, arr, ng, len(x),
             codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
             coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(coeffs), self.hamiltonian.convention,
             gm.shift, mode, 1 if isinstance(opt, Adam) else 0, opt.learning_rate, opt.max_iterations, opt.tolerance,
             x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
             final.handle, ctypes.byref(res))
        return VQEResult(res.optimal_value, x.tolist(), hist[:res.num_iterations].tolist(), final)

    def run_to_target(self, initial_parameters, target_energy):
        """VQE::run with Adam that also stops at the first energy at or below
        target_energy (lq_vqe_run_to_target; no reference counterpart).
        Returns (VQEResult without final state, reached)."""
        arr, ng = _template_args(self.ansatz)
        codes, coeffs = _obs_args(self.hamiltonian)
        x = np.array(initial_parameters, dtype=np.float64, copy=True)
        opt = self.optimizer
        if not isinstance(opt, Adam):
            raise ValueError("run_to_target drives Adam")
        hist = np.zeros(max(1, opt.max_iterations))
        res = LqOptResult()
        gm = self.gradient_method
        mode = getattr(gm, "mode", abi.GRAD_BATCHED if getattr(gm, "batched", False) else abi.GRAD_SEQUENTIAL)
        dtype = getattr(gm, "dtype", C128)
        reached = ctypes.c_int()
        call("lq_vqe_run_to_target", self.ansatz.num_qubits(), dtype, arr, ng, len(x),
             codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
             coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(coeffs), self.hamiltonian.convention,
             gm.shift, mode, opt.learning_rate, opt.max_iterations, opt.tolerance, float(target_energy),
             x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
             ctypes.byref(res), ctypes.byref(reached))
        return VQEResult(res.optimal_value, x.tolist(), hist[:res.num_iterations].tolist(), None), bool(reached.value)

    def run_random(self, seed=None):
        rng = np.random.default_rng(seed)
        return self.run(rng.uniform(0.0, 2.0 * math.pi, self.ansatz.num_parameters()))
