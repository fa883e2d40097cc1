"""Batched evaluation on the GPU: many circuits (of uneven sizes) with their
own Pauli sums in one call (lq_batch_expectations)."""
import ctypes

import numpy as np

from . import abi
from ._lib import call
from .states import C128


class BatchArgs:
    """The C-ABI arguments of a batch, marshalled once: per circuit its qubit
    count, gate array, Pauli code matrix and coefficients.  Passing the same
    BatchArgs to several batch_expectations calls skips rebuilding them."""

    def __init__(self, items):
        count = len(items)
        self.count = count
        self.nq = (ctypes.c_int * count)(*[int(it[0]) for it in items])
        self.ng = (ctypes.c_int * count)(*[len(it[1]) for it in items])
        self.nt = (ctypes.c_int * count)(*[len(it[3]) for it in items])
        self._keep = []
        self.circ = (ctypes.POINTER(abi.LqGate) * count)()
        self.pa = (ctypes.POINTER(ctypes.c_uint8) * count)()
        self.co = (ctypes.POINTER(ctypes.c_double) * count)()
        for c, (n, gates, codes, coeffs) in enumerate(items):
            arr = abi.gate_array(gates)
            codes = np.ascontiguousarray(codes, dtype=np.uint8)
            coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
            self._keep += [arr, codes, coeffs]
            self.circ[c] = ctypes.cast(arr, ctypes.POINTER(abi.LqGate))
            self.pa[c] = codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
            self.co[c] = coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def batch_expectations(items, dtype=C128, convention=abi.PAULI_REFERENCE):
    """items: sequence of (num_qubits, gate descriptors, pauli codes, coeffs),
    or a BatchArgs made from one.  Returns one f64 energy per circuit."""
    a = items if isinstance(items, BatchArgs) else BatchArgs(items)
    out = np.zeros(a.count)
    call("lq_batch_expectations", a.count, a.nq, a.circ, a.ng, a.pa, a.co, a.nt, convention, dtype,
         out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out
