"""C-ABI checks that need no GPU: the library loads, exports every function
include/logosq_hip.h declares, and rejects bad arguments before touching a
device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(lq_\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if not n.endswith("_fn")))


def test_header_declares_functions():
    names = declared_functions()
    assert "lq_state_create" in names and "lq_apply_gate" in names and "lq_expect_pauli_sum" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    import logosq_amd
    lib = logosq_amd.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    from logosq_amd import _lib
    missing = [n for n in declared_functions() if n not in _lib.SIGNATURES and n != "lq_last_error"]
    assert not missing, missing


def test_argument_errors_without_device():
    import logosq_amd
    lib = logosq_amd.load()
    h = ctypes.c_void_p()
    assert lib.lq_state_create(0, 0, ctypes.byref(h)) == logosq_amd.abi.LQ_ERR_INVALID_ARGUMENT
    assert lib.lq_state_create(35, 0, ctypes.byref(h)) == logosq_amd.abi.LQ_ERR_INVALID_ARGUMENT
    assert lib.lq_state_create(4, 7, ctypes.byref(h)) == logosq_amd.abi.LQ_ERR_INVALID_ARGUMENT
    assert b"num_qubits" in lib.lq_last_error() or b"dtype" in lib.lq_last_error()
    assert lib.lq_apply_gate(None, None) == logosq_amd.abi.LQ_ERR_INVALID_ARGUMENT
    assert lib.lq_abi_version() == 1


def test_gate_struct_layout_matches_header():
    from logosq_amd import abi
    assert ctypes.sizeof(abi.LqGate) == 4 + 12 + 24 + 12 + 4 + 256
    assert abi.LqGate.matrix.offset == 56
