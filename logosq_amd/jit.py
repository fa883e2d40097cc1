"""Run-time specialisation of the fused engine's tile passes (lq_set_jit):
each pass structure gets its own kernel, compiled with hipRTC, with gate
kinds, positions and control masks fixed at compile time."""
import ctypes

from . import abi
from ._lib import call

OFF, REPEAT, ALWAYS = 0, 1, 2


def set_jit(mode):
    """OFF: generic tile kernel only; REPEAT (default): compile a pass
    structure the second time a plan holds it (passes of >= 1024 tiles);
    ALWAYS: compile every structure at first sight."""
    call("lq_set_jit", int(mode))


def jit_stats():
    """dict(compiled, failed, launches, compile_seconds) for this process."""
    c, f, l = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
    s = ctypes.c_double()
    call("lq_jit_stats", ctypes.byref(c), ctypes.byref(f), ctypes.byref(l), ctypes.byref(s))
    return {"compiled": c.value, "failed": f.value, "launches": l.value, "compile_seconds": s.value}


def compile_plan(num_qubits, gates, dtype=abi.LQ_C128):
    """Compile the specialised kernels of a circuit's fused plan ahead of
    time (no device needed); returns the number of kernels ready."""
    arr = abi.gate_array(gates)
    k = ctypes.c_int()
    call("lq_jit_compile_plan", int(num_qubits), int(dtype), arr, len(gates), ctypes.byref(k))
    return k.value
