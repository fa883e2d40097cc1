"""Algorithms on the GPU path: QFT circuits (src/algorithms/qft.rs:10-46),
XYZ Heisenberg Trotter evolution and energy
(src/algorithms/xyz_heisenberg.rs:8-162), the seeded brickwork circuit
modelled on rqc::generate_supremacy_circuit (src/algorithms/rqc.rs:236-268),
and the documented input generators."""
import math

import numpy as np

from .circuits import Circuit

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): the seeded generator behind every
    synthetic input; uniforms take the top 53 bits."""

    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def normal(self):
        # Box-Muller, one value per call
        u1 = self.uniform()
        u2 = self.uniform()
        if u1 <= 0.0:
            u1 = 1.0 / 9007199254740992.0
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def normal_state_vector(n, seed):
    """psi_k = N(0,1) + i N(0,1) i.i.d. (unnormalised; State::new normalises)."""
    r = SplitMix64(seed)
    dim = 1 << n
    v = np.empty(dim, dtype=np.complex128)
    for k in range(dim):
        re = r.normal()
        im = r.normal()
        v[k] = complex(re, im)
    return v


def qft_circuit(num_qubits):
    """qft::create_circuit: dense-builder H, CPhaseGate(i, j, pi/2^(j-i)), swaps."""
    c = Circuit(num_qubits).with_name("Quantum Fourier Transform")
    for i in range(num_qubits):
        c.h(i)
        for j in range(i + 1, num_qubits):
            c.controlled_phase(i, j, math.pi / float(1 << (j - i)))
    for i in range(num_qubits // 2):
        c.swap(i, num_qubits - i - 1)
    return c


def inverse_qft_circuit(num_qubits):
    """qft::create_inverse_circuit (qft.rs:30-46)."""
    c = Circuit(num_qubits).with_name("Inverse Quantum Fourier Transform")
    for i in range(num_qubits // 2):
        c.swap(i, num_qubits - i - 1)
    for i in reversed(range(num_qubits)):
        for j in range(i + 1, num_qubits):
            c.controlled_phase(i, j, -math.pi / float(1 << (j - i)))
        c.h(i)
    return c


class HeisenbergParameters:
    """HeisenbergParameters (src/algorithms/xyz_heisenberg.rs:8-28), with the
    reference's defaults."""

    def __init__(self, jx=1.0, jy=1.0, jz=1.0, external_field=0.0, time_steps=100, dt=0.01):
        self.jx, self.jy, self.jz = float(jx), float(jy), float(jz)
        self.external_field = float(external_field)
        self.time_steps = int(time_steps)
        self.dt = float(dt)


def xyz_heisenberg_circuit(num_qubits, params):
    """xyz_heisenberg::create_circuit (xyz_heisenberg.rs:31-61): per Trotter
    step RXX(i, i+1, 2 Jx dt) along the open chain, then RYY(2 Jy dt), then
    RZZ(2 Jz dt), then RZ(i, h dt) on every qubit when h != 0.  RXX/RYY keep
    the reference's orientation rule (qubit1 < qubit2 here, so their 01<->10
    block is left alone, src/gates/parameterized.rs:563-575)."""
    c = Circuit(num_qubits).with_name("XYZ-Heisenberg Model")
    for _ in range(params.time_steps):
        for i in range(num_qubits - 1):
            c.rxx(i, i + 1, 2.0 * (params.jx * params.dt))
        for i in range(num_qubits - 1):
            c.ryy(i, i + 1, 2.0 * (params.jy * params.dt))
        for i in range(num_qubits - 1):
            c.rzz(i, i + 1, 2.0 * (params.jz * params.dt))
        if params.external_field != 0.0:
            for i in range(num_qubits):
                c.rz(i, params.external_field * params.dt)
    return c


def xyz_heisenberg_simulate(state, params):
    """xyz_heisenberg::simulate (xyz_heisenberg.rs:64-67)."""
    xyz_heisenberg_circuit(state.num_qubits(), params).execute(state)


def xyz_heisenberg_observable(num_qubits, params):
    """The open-chain Hamiltonian of calculate_energy_efficient
    (xyz_heisenberg.rs:143-162): +Jx XX + Jy YY + Jz ZZ on bonds (i, i+1),
    then h Z_i when h != 0, as a Pauli sum.  YY carries two Y factors, so
    the reference's Y phase convention (expectation_yy, 190-227) gives the
    physical value."""
    from .optimization import Pauli, PauliObservable, PauliTerm
    o = PauliObservable(num_qubits)
    for i in range(num_qubits - 1):
        for pa, c in ((Pauli.X, params.jx), (Pauli.Y, params.jy), (Pauli.Z, params.jz)):
            p = [Pauli.I] * num_qubits
            p[i] = p[i + 1] = pa
            o.add_term(PauliTerm(c, p))
    if params.external_field != 0.0:
        for i in range(num_qubits):
            p = [Pauli.I] * num_qubits
            p[i] = Pauli.Z
            o.add_term(PauliTerm(params.external_field, p))
    return o


def xyz_energy_efficient(state, params):
    """calculate_energy_efficient (xyz_heisenberg.rs:143-162) on the device:
    one Pauli-sum expectation with f64 accumulation."""
    return xyz_heisenberg_observable(state.num_qubits(), params).expectation(state)


def brickwork_circuit(num_qubits, depth, seed):
    """Random brickwork: per layer one of RX/RY/RZ (uniform) with angle
    U[0, 2pi) on every qubit, then CNOT(q, q+1) for q = layer%2, layer%2+2, ..."""
    r = SplitMix64(seed)
    c = Circuit(num_qubits).with_name("brickwork(d=%d)" % depth)
    for layer in range(depth):
        for q in range(num_qubits):
            kind = min(2, int(r.uniform() * 3.0))
            theta = 2.0 * math.pi * r.uniform()
            (c.rx, c.ry, c.rz)[kind](q, theta)
        for q in range(layer % 2, num_qubits - 1, 2):
            c.cnot(q, q + 1)
    return c


def ry_rz_ring_template(num_qubits, layers):
    """Gate template of the benchmark ansatz: per layer RY(q), RZ(q) on every
    qubit then CZ(q, q+1 mod n); parameter order layer -> qubit -> (RY, RZ)."""
    from . import abi
    t, k = [], 0
    for _ in range(layers):
        for q in range(num_qubits):
            t.append(abi.make_gate(abi.RY, [q], [0.0], [k]))
            t.append(abi.make_gate(abi.RZ, [q], [0.0], [k + 1]))
            k += 2
        for q in range(num_qubits):
            t.append(abi.make_gate(abi.CZ, [q, (q + 1) % num_qubits]))
    return t, k


def bind(template, params):
    """Descriptors with param slots replaced by the given angles."""
    from . import abi
    out = []
    for g in template:
        h = abi.LqGate()
        ctypes_copy(h, g)
        for k in range(3):
            if h.param[k] >= 0:
                h.theta[k] = float(params[h.param[k]])
                h.param[k] = -1
        out.append(h)
    return out


def ctypes_copy(dst, src):
    import ctypes
    ctypes.memmove(ctypes.addressof(dst), ctypes.addressof(src), ctypes.sizeof(src))


def heisenberg_codes(num_qubits, jx, jy, jz):
    """Pauli codes / coefficients of the periodic XYZ ring (3n terms)."""
    codes, coeffs = [], []
    for i in range(num_qubits):
        j = (i + 1) % num_qubits
        for p, c in ((1, jx), (2, jy), (3, jz)):
            row = [0] * num_qubits
            row[i] = row[j] = p
            codes.append(row)
            coeffs.append(c)
    return np.array(codes, dtype=np.uint8), np.array(coeffs, dtype=np.float64)


def ps_gradient_workload(num_qubits=16, layers=6, seed=3):
    """Batched parameter-shift workload: couplings (Jx, Jy, Jz) ~ U[-1,1]^3
    shared by every bond, params ~ U[0, 2pi)^(2nL)."""
    r = SplitMix64(seed)
    j = [2.0 * r.uniform() - 1.0 for _ in range(3)]
    tmpl, P = ry_rz_ring_template(num_qubits, layers)
    params = np.array([2.0 * math.pi * r.uniform() for _ in range(P)])
    codes, coeffs = heisenberg_codes(num_qubits, *j)
    return tmpl, params, codes, coeffs, j


def uneven_batch_workload(count=512, nmin=10, nmax=22, layers=4, terms=32, seed=4):
    """Uneven batch: circuit c has n_c ~ U{nmin..nmax} qubits, the ring ansatz
    with L layers and params ~ U[0, 2pi), and a random `terms`-term Pauli sum
    (each qubit's Pauli uniform over I, X, Y, Z; coefficients N(0,1))."""
    r = SplitMix64(seed)
    out = []
    for _ in range(count):
        n = nmin + min(nmax - nmin, int(r.uniform() * (nmax - nmin + 1)))
        tmpl, P = ry_rz_ring_template(n, layers)
        params = [2.0 * math.pi * r.uniform() for _ in range(P)]
        codes = np.array([[min(3, int(r.uniform() * 4)) for _ in range(n)] for _ in range(terms)], dtype=np.uint8)
        coeffs = np.array([r.normal() for _ in range(terms)])
        out.append((n, bind(tmpl, params), codes, coeffs))
    return out
