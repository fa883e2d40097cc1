"""Batched evaluation on the GPU: many circuits (of uneven sizes) with their
own Pauli sums in one call (lq_batch_expectations)."""
import ctypes

import numpy as np

from . import abi
from ._lib import call
from .states import C128


def batch_expectations(items, dtype=C128, convention=abi.PAULI_REFERENCE):
    """items: sequence of (num_qubits, gate descriptors, pauli codes, coeffs).
    Returns one f64 energy per circuit."""
    count = len(items)
    nq = (ctypes.c_int * count)(*[int(it[0]) for it in items])
    ng = (ctypes.c_int * count)(*[len(it[1]) for it in items])
    nt = (ctypes.c_int * count)(*[len(it[3]) for it in items])
    keep = []
    circ = (ctypes.POINTER(abi.LqGate) * count)()
    pa = (ctypes.POINTER(ctypes.c_uint8) * count)()
    co = (ctypes.POINTER(ctypes.c_double) * count)()
    for c, (n, gates, codes, coeffs) in enumerate(items):
        arr = abi.gate_array(gates)
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        keep += [arr, codes, coeffs]
        circ[c] = ctypes.cast(arr, ctypes.POINTER(abi.LqGate))
        pa[c] = codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        co[c] = coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    out = np.zeros(count)
    call("lq_batch_expectations", count, nq, circ, ng, pa, co, nt, convention, dtype,
         out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out
