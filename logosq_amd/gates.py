"""Gates: mirrors of the reference's gate structs (src/gates/*.rs).

Each struct's `apply(state)` runs one HIP kernel through lq_apply_gate; the
`desc()` method returns the lq_gate descriptor a Circuit records.
"""
import ctypes
import math

import numpy as np

from . import abi
from ._lib import call


class Gate:
    """trait Gate (src/gates/mod.rs:26-35)."""

    def desc(self):
        raise NotImplementedError

    def apply(self, state):
        g = self.desc()
        call("lq_apply_gate", state.handle, ctypes.byref(g))


class _G(Gate):
    kind = None
    nq = 1

    def __init__(self, *args):
        # (qubits..., [angles...], num_qubits)
        self.qubits = list(args[: self.nq])
        self.angles = list(args[self.nq: -1])
        self.num_qubits = args[-1]

    def desc(self):
        return abi.make_gate(self.kind, self.qubits, self.angles)


def _mk(name, kind, nq):
    return type(name, (_G,), {"kind": kind, "nq": nq})


XGate = _mk("XGate", abi.X, 1)               # single_qubit.rs:9-31
YGate = _mk("YGate", abi.Y, 1)               # single_qubit.rs:34-59 (applies -Y)
ZGate = _mk("ZGate", abi.Z, 1)               # single_qubit.rs:62-82
HGate = _mk("HGate", abi.H, 1)               # single_qubit.rs:85-113
SGate = _mk("SGate", abi.S, 1)               # single_qubit.rs:116-137
TGate = _mk("TGate", abi.T, 1)               # single_qubit.rs:140-161
RXGate = _mk("RXGate", abi.RX, 1)            # parameterized.rs:31-78
RYGate = _mk("RYGate", abi.RY, 1)            # parameterized.rs:84-131
RZGate = _mk("RZGate", abi.RZ, 1)            # parameterized.rs:137-176
PhaseGate = _mk("PhaseGate", abi.PHASE, 1)   # parameterized.rs:182-218
U3Gate = _mk("U3Gate", abi.U3, 1)            # parameterized.rs:228-282
CNOTGate = _mk("CNOTGate", abi.CNOT, 2)      # two_qubit.rs:6-31
SWAPGate = _mk("SWAPGate", abi.SWAP, 2)      # two_qubit.rs:34-62
CZGate = _mk("CZGate", abi.CZ, 2)            # two_qubit.rs:65-90
CPhaseGate = _mk("CPhaseGate", abi.CPHASE, 2)  # two_qubit.rs:93-136
CRXGate = _mk("CRXGate", abi.CRX, 2)         # parameterized.rs:292-343
CRYGate = _mk("CRYGate", abi.CRY, 2)         # parameterized.rs:349-399
CRZGate = _mk("CRZGate", abi.CRZ, 2)         # parameterized.rs:405-451
CPhaseParamGate = _mk("CPhaseParamGate", abi.CP, 2)  # parameterized.rs:457-499
RXXGate = _mk("RXXGate", abi.RXX, 2)         # parameterized.rs:509-578
RYYGate = _mk("RYYGate", abi.RYY, 2)         # parameterized.rs:585-653
RZZGate = _mk("RZZGate", abi.RZZ, 2)         # parameterized.rs:660-709
ToffoliGate = _mk("ToffoliGate", abi.TOFFOLI, 3)  # three_qubit.rs:6-35


# convenience constructors (single_qubit.rs:164-191, two_qubit.rs:139-170, ...)
def x(q, n): return XGate(q, n)
def y(q, n): return YGate(q, n)
def z(q, n): return ZGate(q, n)
def h(q, n): return HGate(q, n)
def s(q, n): return SGate(q, n)
def t(q, n): return TGate(q, n)
def rx(q, th, n): return RXGate(q, th, n)
def ry(q, th, n): return RYGate(q, th, n)
def rz(q, th, n): return RZGate(q, th, n)
def phase(q, th, n): return PhaseGate(q, th, n)
def u3(q, th, ph, la, n): return U3Gate(q, th, ph, la, n)
def cnot(c, tq, n): return CNOTGate(c, tq, n)
def swap(a, b, n): return SWAPGate(a, b, n)
def cz(c, tq, n): return CZGate(c, tq, n)
def cphase(c, tq, ang, n): return CPhaseGate(c, tq, ang, n)
def crx(c, tq, th, n): return CRXGate(c, tq, th, n)
def cry(c, tq, th, n): return CRYGate(c, tq, th, n)
def crz(c, tq, th, n): return CRZGate(c, tq, th, n)
def cphase_param(c, tq, th, n): return CPhaseParamGate(c, tq, th, n)
def rxx(a, b, th, n): return RXXGate(a, b, th, n)
def ryy(a, b, th, n): return RYYGate(a, b, th, n)
def rzz(a, b, th, n): return RZZGate(a, b, th, n)
def toffoli(c1, c2, tq, n): return ToffoliGate(c1, c2, tq, n)


class MatrixGate(Gate):
    """MatrixGate (src/gates/mod.rs:56-93): dense matvec then normalize."""

    def __init__(self, matrix):
        self.matrix = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128))

    def apply(self, state):
        dim = self.matrix.shape[1]
        if self.matrix.shape[0] != self.matrix.shape[1]:
            raise ValueError("MatrixGate needs a square matrix")
        p = self.matrix.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        call("lq_apply_dense", state.handle, p, dim)


# ---- dense matrices (src/gates/matrix.rs:14-261) ----
_R2 = 1.0 / math.sqrt(2.0)


def x_gate(): return MatrixGate([[0, 1], [1, 0]])
def y_gate(): return MatrixGate([[0, -1j], [1j, 0]])
def z_gate(): return MatrixGate([[1, 0], [0, -1]])
def h_gate(): return MatrixGate([[_R2, _R2], [_R2, -_R2]])
def s_gate(): return MatrixGate([[1, 0], [0, 1j]])
def t_gate(): return MatrixGate([[1, 0], [0, complex(_R2, _R2)]])


def rx_gate(theta):
    c, s_ = math.cos(theta / 2.0), math.sin(theta / 2.0)
    return MatrixGate([[c, -1j * s_], [-1j * s_, c]])


def ry_gate(theta):
    c, s_ = math.cos(theta / 2.0), math.sin(theta / 2.0)
    return MatrixGate([[c, -s_], [s_, c]])


def rz_gate(theta):
    return MatrixGate([[complex(math.cos(theta / 2.0), -math.sin(theta / 2.0)), 0],
                       [0, complex(math.cos(theta / 2.0), math.sin(theta / 2.0))]])


def cnot_gate():
    return MatrixGate([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]])


def swap_gate():
    return MatrixGate([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])


def cz_gate():
    return MatrixGate(np.diag([1, 1, 1, -1]))


def toffoli_gate():
    m = np.eye(8, dtype=np.complex128)
    m[6, 6] = m[7, 7] = 0
    m[6, 7] = m[7, 6] = 1
    return MatrixGate(m)
