"""ctypes mirror of include/logosq_hip.h (types and constants only; loading
the library lives in logosq_amd._lib)."""
import ctypes

LQ_OK = 0
LQ_ERR_INVALID_QUBIT_INDEX = 1
LQ_ERR_INVALID_STATE_DIMENSION = 2
LQ_ERR_DIMENSION_MISMATCH = 3
LQ_ERR_INVALID_GATE_MATRIX = 4
LQ_ERR_PARAMETER_COUNT_MISMATCH = 5
LQ_ERR_CIRCUIT_QUBIT_MISMATCH = 6
LQ_ERR_INVALID_ARGUMENT = 20
LQ_ERR_OUT_OF_MEMORY = 21
LQ_ERR_DEVICE = 22
LQ_ERR_UNSUPPORTED = 23

LQ_C128 = 0
LQ_C64 = 1

X, Y, Z, H, S, T = 1, 2, 3, 4, 5, 6
RX, RY, RZ, PHASE, U3 = 10, 11, 12, 13, 14
CNOT, SWAP, CZ, CPHASE = 20, 21, 22, 23
CRX, CRY, CRZ, CP = 30, 31, 32, 33
RXX, RYY, RZZ = 40, 41, 42
TOFFOLI = 50
MATRIX1, MATRIX2 = 60, 61

GATE_FLAG_RENORMALIZE = 1

PAULI_I, PAULI_X, PAULI_Y, PAULI_Z = 0, 1, 2, 3
PAULI_REFERENCE, PAULI_PHYSICAL = 0, 1
EXPECT_TILES_OFF, EXPECT_TILES_AUTO, EXPECT_TILES_ON = 0, 1, 2

EXEC_PER_GATE, EXEC_FUSED = 0, 1
UNITS_PAIR_HI, UNITS_PAIR_LO, UNITS_DIAG, UNITS_SWAP, UNITS_QUAD, UNITS_QUAD_LO, UNITS_ALL = 1, 2, 4, 8, 16, 32, 63
GRAD_SEQUENTIAL, GRAD_BATCHED, GRAD_ADJOINT = 0, 1, 2

ARITY = {X: 1, Y: 1, Z: 1, H: 1, S: 1, T: 1, RX: 1, RY: 1, RZ: 1, PHASE: 1, U3: 1, MATRIX1: 1,
         CNOT: 2, SWAP: 2, CZ: 2, CPHASE: 2, CRX: 2, CRY: 2, CRZ: 2, CP: 2, RXX: 2, RYY: 2, RZZ: 2,
         MATRIX2: 2, TOFFOLI: 3}


class LqGate(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("qubits", ctypes.c_int32 * 3),
        ("theta", ctypes.c_double * 3),
        ("param", ctypes.c_int32 * 3),
        ("flags", ctypes.c_int32),
        ("matrix", ctypes.c_double * 32),
    ]


def make_gate(kind, qubits, theta=(), param=(), flags=0, matrix=None):
    """Build one LqGate.  `matrix` is a sequence of complex numbers (row-major)."""
    g = LqGate()
    g.kind = kind
    qs = list(qubits) + [-1] * (3 - len(qubits))
    for i in range(3):
        g.qubits[i] = qs[i]
    th = list(theta) + [0.0] * (3 - len(theta))
    for i in range(3):
        g.theta[i] = float(th[i])
    pr = list(param) + [-1] * (3 - len(param))
    for i in range(3):
        g.param[i] = int(pr[i])
    g.flags = flags
    if matrix is not None:
        flat = []
        for z in matrix:
            z = complex(z)
            flat += [z.real, z.imag]
        for i, v in enumerate(flat):
            g.matrix[i] = v
    return g


def gate_array(gates):
    arr = (LqGate * max(1, len(gates)))()
    for i, g in enumerate(gates):
        arr[i] = g
    return arr
