"""Loads the in-tree HIP library liblogosq_hip.so and declares its C ABI.

There is no CPU fallback: if the library is missing or no GPU is present the
calls raise.  The library is found next to this file (built in tree by
logosq_amd/build.py) or at $LOGOSQ_HIP_LIB.
"""
import ctypes
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LOGOSQ_HIP_LIB", os.path.join(_HERE, "liblogosq_hip.so"))

_lib = None

c_int = ctypes.c_int
c_u64 = ctypes.c_uint64
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
P_double = ctypes.POINTER(ctypes.c_double)
P_float = ctypes.POINTER(ctypes.c_float)
P_u8 = ctypes.POINTER(ctypes.c_uint8)
P_int = ctypes.POINTER(ctypes.c_int)
P_gate = ctypes.POINTER(abi.LqGate)
OBJ_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_void_p)
GRAD_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                           ctypes.c_void_p)

# name -> argtypes (all functions return int status except lq_last_error)
SIGNATURES = {
    "lq_abi_version": [],
    "lq_device_info": [ctypes.c_char_p, ctypes.c_size_t, P_int, ctypes.POINTER(ctypes.c_size_t),
                       ctypes.POINTER(ctypes.c_size_t)],
    "lq_synchronize": [],
    "lq_event_record": [c_int],
    "lq_event_elapsed_ms": [c_int, c_int, ctypes.POINTER(ctypes.c_float)],
    "lq_state_create": [c_int, c_int, ctypes.POINTER(c_void_p)],
    "lq_state_destroy": [c_void_p],
    "lq_state_num_qubits": [c_void_p, P_int],
    "lq_state_dtype": [c_void_p, P_int],
    "lq_state_set_basis": [c_void_p, c_u64],
    "lq_state_set_plus": [c_void_p],
    "lq_state_upload": [c_void_p, P_double, c_u64, c_int],
    "lq_state_download": [c_void_p, P_double, c_u64],
    "lq_state_copy": [c_void_p, c_void_p],
    "lq_state_normalize": [c_void_p, P_double],
    "lq_state_norm_sqr": [c_void_p, P_double],
    "lq_state_inner_product": [c_void_p, c_void_p, P_double],
    "lq_state_probabilities": [c_void_p, P_double, c_u64],
    "lq_bit_probabilities": [c_void_p, P_double],
    "lq_measure_qubit": [c_void_p, c_int, c_double, P_int],
    "lq_sample_shots": [c_void_p, c_u64, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64)],
    "lq_noise_amplitude_damping": [c_void_p, P_int, c_int, c_double],
    "lq_noise_phase_damping": [c_void_p, P_int, c_int, c_double],
    "lq_noise_depolarizing": [c_void_p, P_int, c_int, c_double, ctypes.POINTER(ctypes.c_uint64), P_int],
    "lq_apply_gate": [c_void_p, P_gate],
    "lq_apply_1q": [c_void_p, c_int, P_double],
    "lq_apply_ctrl_1q": [c_void_p, P_int, c_int, c_int, P_double],
    "lq_apply_2q": [c_void_p, c_int, c_int, P_double],
    "lq_apply_dense": [c_void_p, P_double, c_u64],
    "lq_execute": [c_void_p, P_gate, c_int, c_int],
    "lq_set_fused_config": [c_int, c_int, c_int],
    "lq_set_gate_kernel_config": [c_int, c_int],
    "lq_set_gate_unit_order": [c_int],
    "lq_autotune": [c_int, c_int, ctypes.c_char_p, ctypes.c_size_t],
    "lq_plan_fused": [c_int, c_int, P_gate, c_int, c_int, c_int, P_int, P_int, P_int],
    "lq_expect_pauli_sum": [c_void_p, P_u8, P_double, c_int, c_int, P_double],
    "lq_expect_pauli_sum_noalloc": [c_void_p, P_u8, P_double, c_int, c_int, P_double],
    "lq_param_shift_gradient": [c_int, c_int, P_gate, c_int, P_double, c_int, c_double, P_u8, P_double, c_int,
                                c_int, c_int, P_double, P_double],
    "lq_adjoint_gradient": [c_int, c_int, P_gate, c_int, P_double, c_int, P_u8, P_double, c_int, c_int, P_double,
                            P_double],
    "lq_adjoint_gradient_batch": [c_int, c_int, P_gate, c_int, c_int, P_double, c_int, P_u8, P_double, c_int, c_int,
                                  P_double, P_double],
    "lq_vqe_run_to_target": [c_int, c_int, P_gate, c_int, c_int, P_u8, P_double, c_int, c_int, c_double, c_int,
                             c_double, c_int, c_double, c_double, P_double, P_double, ctypes.c_void_p, P_int],
    "lq_vqe_adam_batch": [c_int, c_int, P_gate, c_int, c_int, P_u8, P_double, c_int, c_int, c_int, c_double, c_int,
                          c_double, P_double, P_double, P_double, P_int, P_int],
    "lq_set_batch_graphs": [c_int],
    "lq_batch_graph_stats": [ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)],
    "lq_set_jit": [c_int],
    "lq_set_expect_tiles": [c_int],
    "lq_expect_plan": [c_int, c_int, P_u8, P_double, c_int, c_int, P_int, P_int, P_int],
    "lq_expect_jit_compile": [c_int, c_int, P_u8, P_double, c_int, c_int, P_int],
    "lq_expect_jit_describe": [c_int, c_int, P_u8, P_double, c_int, c_int, ctypes.c_char_p, ctypes.c_size_t,
                               ctypes.POINTER(ctypes.c_size_t)],
    "lq_jit_stats": [ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                     ctypes.POINTER(ctypes.c_longlong), P_double],
    "lq_jit_compile_plan": [c_int, c_int, P_gate, c_int, P_int],
    "lq_batch_expectations": [c_int, P_int, ctypes.POINTER(P_gate), P_int, ctypes.POINTER(P_u8),
                              ctypes.POINTER(P_double), P_int, c_int, c_int, P_double],
    "lq_gd_minimize": [c_double, c_int, c_double, OBJ_FN, GRAD_FN, c_void_p, P_double, c_int, P_double, c_void_p],
    "lq_adam_minimize": [c_double, c_double, c_double, c_double, c_int, c_double, OBJ_FN, GRAD_FN, c_void_p,
                         P_double, c_int, P_double, c_void_p],
    "lq_vqe_run": [c_int, c_int, P_gate, c_int, c_int, P_u8, P_double, c_int, c_int, c_double, c_int, c_int,
                   c_double, c_int, c_double, P_double, P_double, c_void_p, c_void_p],
    "lq_mlp_create": [c_int, c_int, c_int, c_u64, ctypes.POINTER(c_void_p)],
    "lq_mlp_destroy": [c_void_p],
    "lq_mlp_num_params": [c_void_p, ctypes.POINTER(ctypes.c_int64)],
    "lq_mlp_get_params": [c_void_p, P_float, ctypes.c_int64],
    "lq_mlp_set_params": [c_void_p, P_float, ctypes.c_int64],
    "lq_mlp_forward": [c_void_p, P_float, c_int, P_float],
    "lq_mlp_gradients": [c_void_p, P_float, P_float, c_int, P_float, P_float],
    "lq_mlp_train": [c_void_p, P_float, P_float, c_int, c_int, c_int, ctypes.c_float, c_u64, P_float],
}


class LogosQError(RuntimeError):
    """Mirror of LogosQError (src/error.rs:9-62); `.code` is the lq_status."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class InvalidQubitIndex(LogosQError):
    pass


class InvalidStateDimension(LogosQError):
    pass


class DimensionMismatch(LogosQError):
    pass


class ParameterCountMismatch(LogosQError):
    pass


class CircuitQubitMismatch(LogosQError):
    pass


class DeviceError(LogosQError):
    pass


_ERR_CLASS = {
    abi.LQ_ERR_INVALID_QUBIT_INDEX: InvalidQubitIndex,
    abi.LQ_ERR_INVALID_STATE_DIMENSION: InvalidStateDimension,
    abi.LQ_ERR_DIMENSION_MISMATCH: DimensionMismatch,
    abi.LQ_ERR_PARAMETER_COUNT_MISMATCH: ParameterCountMismatch,
    abi.LQ_ERR_CIRCUIT_QUBIT_MISMATCH: CircuitQubitMismatch,
    abi.LQ_ERR_DEVICE: DeviceError,
    abi.LQ_ERR_OUT_OF_MEMORY: DeviceError,
}


def load(path=None):
    """Load (once) and return the ctypes library.  Raises OSError if absent."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError("liblogosq_hip.so not found at %s: build it with `python logosq_amd/build.py`" % p)
    lib = ctypes.CDLL(p)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = c_int
    lib.lq_last_error.argtypes = []
    lib.lq_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def check(rc):
    if rc != abi.LQ_OK:
        msg = _lib.lq_last_error().decode(errors="replace") if _lib is not None else ""
        raise _ERR_CLASS.get(rc, LogosQError)(rc, msg or ("lq status %d" % rc))
    return rc


def call(name, *args):
    lib = load()
    return check(getattr(lib, name)(*args))
