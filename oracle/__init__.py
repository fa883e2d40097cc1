"""Python access to the CPU oracle (oracle/lq_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product path never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

from logosq_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
# LQ_ORACLE_LIB selects another build (the sanitized one: make -C oracle sanitize)
LIB = os.environ.get("LQ_ORACLE_LIB", os.path.join(HERE, "_build", "liblq_oracle.so"))

_lib = None


def build():
    r = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB


class OptResult(ctypes.Structure):
    _fields_ = [("optimal_value", ctypes.c_double), ("num_iterations", ctypes.c_int),
                ("converged", ctypes.c_int)]


OBJ = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_void_p)
GRD = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                       ctypes.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) and "LQ_ORACLE_LIB" not in os.environ:
            build()
        L = ctypes.CDLL(LIB)
        d, i, u64, vp = ctypes.c_double, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p
        pd = ctypes.POINTER(ctypes.c_double)
        pg = ctypes.POINTER(abi.LqGate)
        pu8 = ctypes.POINTER(ctypes.c_uint8)
        sig = {
            "or_normalize": (d, [vp, u64]),
            "or_inner_product": (None, [vp, vp, u64, ctypes.c_void_p]),
            "or_apply_gate": (i, [vp, i, pg]),
            "or_execute": (i, [vp, i, pg, i]),
            "or_matrix_gate": (None, [vp, i, vp, u64, vp]),
            "or_kron_single": (None, [vp, i, i, vp]),
            "or_expand_two": (None, [vp, i, i, i, vp]),
            "or_qft_circuit": (i, [i, pg]),
            "or_inverse_qft_circuit": (i, [i, pg]),
            "or_pauli_term_expectation": (d, [vp, i, pu8, d]),
            "or_pauli_sum_expectation": (d, [vp, i, pu8, pd, i]),
            "or_param_shift": (None, [i, pg, i, pd, i, d, pu8, pd, i, pd]),
            "or_param_shift_range": (None, [i, pg, i, pd, i, d, pu8, pd, i, i, i, pd]),
            "or_finite_difference": (None, [i, pg, i, pd, i, d, pu8, pd, i, pd]),
            "or_energy": (d, [i, pg, i, pd, pu8, pd, i]),
            "or_gd_minimize": (OptResult, [d, i, d, OBJ, GRD, vp, pd, i, pd]),
            "or_adam_minimize": (OptResult, [d, d, d, d, i, d, OBJ, GRD, vp, pd, i, pd]),
            "or_vqe_run": (OptResult, [i, pg, i, i, pu8, pd, i, d, i, d, i, d, pd, pd, vp]),
            "or_xyz_energy_efficient": (d, [vp, i, d, d, d, d]),
            "or_now": (d, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _c(v):
    v = np.ascontiguousarray(v, dtype=np.complex128)
    return v, v.ctypes.data_as(ctypes.c_void_p)


def _pd(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def zero_state(n):
    v = np.zeros(1 << n, dtype=np.complex128)
    v[0] = 1.0
    return v


def normalize(v):
    v, p = _c(np.array(v, dtype=np.complex128, copy=True))
    norm = lib().or_normalize(p, len(v))
    return v, norm


def inner_product(a, b):
    """<a|b> (src/states/mod.rs:350-391), summed in index order."""
    a, pa = _c(a)
    b, pb = _c(b)
    out = (ctypes.c_double * 2)()
    lib().or_inner_product(pa, pb, len(a), out)
    return complex(out[0], out[1])


def apply_gate(v, n, g):
    v, p = _c(np.array(v, dtype=np.complex128, copy=True))
    rc = lib().or_apply_gate(p, n, ctypes.byref(g))
    assert rc == 0
    return v


def execute(v, n, gates):
    v, p = _c(np.array(v, dtype=np.complex128, copy=True))
    arr = abi.gate_array(gates)
    rc = lib().or_execute(p, n, arr, len(gates))
    assert rc == 0
    return v


def matrix_gate(v, n, m):
    v, p = _c(np.array(v, dtype=np.complex128, copy=True))
    m = np.ascontiguousarray(m, dtype=np.complex128)
    scratch = np.empty_like(v)
    lib().or_matrix_gate(p, n, m.ctypes.data_as(ctypes.c_void_p), len(v), scratch.ctypes.data_as(ctypes.c_void_p))
    return v


def kron_single(m, n, target):
    m = np.ascontiguousarray(m, dtype=np.complex128)
    full = np.empty((1 << n, 1 << n), dtype=np.complex128)
    lib().or_kron_single(m.ctypes.data_as(ctypes.c_void_p), n, target, full.ctypes.data_as(ctypes.c_void_p))
    return full


def expand_two(m, n, q1, q2):
    m = np.ascontiguousarray(m, dtype=np.complex128)
    full = np.empty((1 << n, 1 << n), dtype=np.complex128)
    lib().or_expand_two(m.ctypes.data_as(ctypes.c_void_p), n, q1, q2, full.ctypes.data_as(ctypes.c_void_p))
    return full


def qft_gates(n, inverse=False):
    f = lib().or_inverse_qft_circuit if inverse else lib().or_qft_circuit
    k = f(n, None)
    arr = (abi.LqGate * k)()
    f(n, arr)
    return list(arr)


def _codes(codes, coeffs):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
    return codes, coeffs


def pauli_sum(v, n, codes, coeffs):
    v, p = _c(v)
    codes, coeffs = _codes(codes, coeffs)
    return lib().or_pauli_sum_expectation(p, n, codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _pd(coeffs),
                                          len(coeffs))


def pauli_term(v, n, paulis, coeff):
    v, p = _c(v)
    paulis = np.ascontiguousarray(paulis, dtype=np.uint8)
    return lib().or_pauli_term_expectation(p, n, paulis.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), coeff)


def energy(n, tmpl, params, codes, coeffs):
    arr = abi.gate_array(tmpl)
    params = np.ascontiguousarray(params, dtype=np.float64)
    codes, coeffs = _codes(codes, coeffs)
    return lib().or_energy(n, arr, len(tmpl), _pd(params), codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                           _pd(coeffs), len(coeffs))


def param_shift(n, tmpl, params, codes, coeffs, shift=np.pi / 2, threads=1):
    """ParameterShift::compute_gradient.  threads > 1 splits the components
    into contiguous ranges evaluated in parallel (the C loop releases the
    GIL and is re-entrant); every component is computed exactly as in the
    serial loop."""
    arr = abi.gate_array(tmpl)
    params = np.ascontiguousarray(params, dtype=np.float64)
    codes, coeffs = _codes(codes, coeffs)
    g = np.zeros(len(params))
    pc = codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    if threads <= 1:
        lib().or_param_shift(n, arr, len(tmpl), _pd(params), len(params), shift, pc, _pd(coeffs), len(coeffs),
                             _pd(g))
        return g
    import concurrent.futures
    P = len(params)
    bounds = [(P * k // threads, P * (k + 1) // threads) for k in range(threads)]

    def part(b):
        lib().or_param_shift_range(n, arr, len(tmpl), _pd(params), P, shift, pc, _pd(coeffs), len(coeffs), b[0], b[1],
                                   _pd(g))

    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        list(ex.map(part, bounds))
    return g


def finite_difference(n, tmpl, params, codes, coeffs, eps=1e-7):
    arr = abi.gate_array(tmpl)
    params = np.ascontiguousarray(params, dtype=np.float64)
    codes, coeffs = _codes(codes, coeffs)
    g = np.zeros(len(params))
    lib().or_finite_difference(n, arr, len(tmpl), _pd(params), len(params), eps,
                               codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _pd(coeffs), len(coeffs), _pd(g))
    return g


def _wrap(f, g):
    def fo(x, dim, _u):
        return float(f([x[i] for i in range(dim)]))

    def go(x, dim, out, _u):
        gr = g([x[i] for i in range(dim)])
        for i in range(dim):
            out[i] = gr[i]

    return OBJ(fo), GRD(go)


def gd_minimize(f, g, x0, lr, max_it, tol=1e-6):
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    hist = np.zeros(max(1, max_it))
    fo, go = _wrap(f, g)
    r = lib().or_gd_minimize(lr, max_it, tol, fo, go, None, _pd(x), len(x), _pd(hist))
    return x, r.optimal_value, hist[: r.num_iterations], bool(r.converged)


def adam_minimize(f, g, x0, lr, max_it, tol=1e-6):
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    hist = np.zeros(max(1, max_it))
    fo, go = _wrap(f, g)
    r = lib().or_adam_minimize(lr, 0.9, 0.999, 1e-8, max_it, tol, fo, go, None, _pd(x), len(x), _pd(hist))
    return x, r.optimal_value, hist[: r.num_iterations], bool(r.converged)


def vqe_run(n, tmpl, num_params, codes, coeffs, x0, optimizer, lr, max_it, tol=1e-6, shift=np.pi / 2):
    arr = abi.gate_array(tmpl)
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    codes, coeffs = _codes(codes, coeffs)
    hist = np.zeros(max(1, max_it))
    final = np.zeros(1 << n, dtype=np.complex128)
    r = lib().or_vqe_run(n, arr, len(tmpl), num_params, codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                         _pd(coeffs), len(coeffs), shift, optimizer, lr, max_it, tol, _pd(x), _pd(hist),
                         final.ctypes.data_as(ctypes.c_void_p))
    return x, r.optimal_value, hist[: r.num_iterations], bool(r.converged), final


def xyz_energy(v, n, jx, jy, jz, field=0.0):
    v, p = _c(v)
    return lib().or_xyz_energy_efficient(p, n, jx, jy, jz, field)


def now():
    return lib().or_now()
