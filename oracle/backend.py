"""CPU restatement of the reference's QuantumStateBackend path for the dense
State: Circuit::execute_on_backend (src/circuits/mod.rs:444-504) with the
matrix extraction helpers (mod.rs:508-590) and State's backend impl
(src/states/mod.rs:394-487, expand_two_qubit_gate 509-549).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker; the product path
never imports it.  Small n only (full 2^n x 2^n matrices).
"""
import numpy as np

import oracle


def full_matrix(n, gate):
    """The gate's full 2^n matrix, probed column by column on basis states
    (mod.rs:456-467)."""
    dim = 1 << n
    m = np.zeros((dim, dim), dtype=np.complex128)
    for col in range(dim):
        v = np.zeros(dim, dtype=np.complex128)
        v[col] = 1.0
        m[:, col] = oracle.apply_gate(v, n, gate)
    return m


def _index(n, assign):
    """Basis index with the given {qubit: bit}; other qubits 0."""
    i = 0
    for q, b in assign.items():
        if b:
            i |= 1 << (n - 1 - q)
    return i


def extract_single(full, qubit, n):
    """extract_single_qubit_matrix (mod.rs:508-545): the block with every
    other qubit 0."""
    g = np.zeros((2, 2), dtype=np.complex128)
    for r in range(2):
        for c in range(2):
            g[r, c] = full[_index(n, {qubit: r}), _index(n, {qubit: c})]
    return g


def extract_two(full, control, target, n):
    """extract_two_qubit_matrix (mod.rs:547-590): sub-index
    (bit(control) << 1) | bit(target), every other qubit 0."""
    assert control != target
    g = np.zeros((4, 4), dtype=np.complex128)
    for r in range(4):
        for c in range(4):
            g[r, c] = full[_index(n, {control: (r >> 1) & 1, target: r & 1}),
                           _index(n, {control: (c >> 1) & 1, target: c & 1})]
    return g


def expand_two_backend(m, n, control, target):
    """expand_two_qubit_gate (src/states/mod.rs:509-549): element (i, j) is
    m[(c_i << 1) | t_i, (c_j << 1) | t_j] when every other qubit agrees."""
    dim = 1 << n
    full = np.zeros((dim, dim), dtype=np.complex128)
    others = [q for q in range(n) if q != control and q != target]
    for i in range(dim):
        for j in range(dim):
            bit = lambda x, q: (x >> (n - 1 - q)) & 1
            if all(bit(i, q) == bit(j, q) for q in others):
                full[i, j] = m[(bit(i, control) << 1) | bit(i, target), (bit(j, control) << 1) | bit(j, target)]
    return full


def apply_full(v, n, full):
    """State::apply_full_matrix (src/states/mod.rs:471-486): dot, normalize."""
    return oracle.matrix_gate(v, n, full)


def execute_on_backend(v, n, gates, op_qubits=None):
    """Circuit::execute_on_backend on the dense State (op_qubits: each
    operation's qubit list, default the gate's own qubits): per operation, the
    probed matrix reduced to the operation's qubits, then the State's
    apply_single_qubit_matrix (kron with identities, mod.rs:399-439),
    apply_two_qubit_matrix (mod.rs:441-457) or apply_three_qubit_matrix
    (the full probed matrix, mod.rs:459-469), each ending in
    apply_full_matrix."""
    from logosq_amd import abi
    v = np.array(v, dtype=np.complex128, copy=True)
    for i, g in enumerate(gates):
        qs = op_qubits[i] if op_qubits is not None else list(g.qubits[:abi.ARITY[g.kind]])
        full = full_matrix(n, g)
        if len(qs) == 1:
            v = apply_full(v, n, oracle.kron_single(extract_single(full, qs[0], n), n, qs[0]))
        elif len(qs) == 2:
            c, t = qs
            v = apply_full(v, n, expand_two_backend(extract_two(full, c, t, n), n, c, t))
        else:
            v = apply_full(v, n, full)
    return v
