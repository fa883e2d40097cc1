"""lq_expect_pauli_sum_noalloc (the expectation with its workspace in a
static device array) against lq_expect_pauli_sum through the C ABI:
two calls in a row, the second on a smaller state and sum, complex128 and
complex64, both within 1e-10 of each other and of the oracle at sizes it
finishes quickly; a sum too large for the 4 MiB array is refused before any
device work and leaves later calls correct."""
import ctypes

import numpy as np
import pytest

import oracle
import logosq_amd as lq
from logosq_amd import abi
from tests import bruteforce as bf
from tests import modes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _args(codes, coeffs):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
    return codes, coeffs, codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), coeffs.ctypes.data_as(
        ctypes.POINTER(ctypes.c_double))


def both(st, codes, coeffs, conv=abi.PAULI_REFERENCE):
    codes, coeffs, pc, pk = _args(codes, coeffs)
    a, b = ctypes.c_double(), ctypes.c_double()
    lq._lib.call("lq_expect_pauli_sum", st.handle, pc, pk, len(coeffs), conv, ctypes.byref(a))
    lq._lib.call("lq_expect_pauli_sum_noalloc", st.handle, pc, pk, len(coeffs), conv, ctypes.byref(b))
    return a.value, b.value


def random_sum(n, terms, rng):
    codes = rng.integers(0, 4, size=(terms, n)).astype(np.uint8)
    return codes, rng.normal(size=terms)


def ring_plus_random(n, rng):
    c1, k1 = lq.algorithms.heisenberg_codes(n, *rng.uniform(-1, 1, 3))
    c2, k2 = random_sum(n, 12, rng)
    return np.vstack([c1, c2]), np.concatenate([k1, k2])


@pytest.mark.parametrize("dtype,tol", [(lq.C128, 1e-10), (lq.C64, 1e-5)])
def test_noalloc_matches_expectation_twice(hip, dtype, tol):
    rng = np.random.default_rng(7700 + dtype)
    try:
        for tiles in (abi.EXPECT_TILES_AUTO, abi.EXPECT_TILES_ON):
            lq._lib.call("lq_set_expect_tiles", tiles)
            for n, make in ((20, ring_plus_random), (9, lambda n, r: random_sum(n, 7, r))):  # second call smaller
                say("noalloc n=%d dtype=%d tiles=%d" % (n, dtype, tiles))
                psi = bf.random_state(n, rng)
                st = lq.State(n, dtype)
                st.set_vector(psi)
                codes, coeffs = make(n, rng)
                a, b = both(st, codes, coeffs)
                assert abs(a - b) <= 1e-12, (a, b)
                want = oracle.pauli_sum(psi, n, codes, coeffs)
                assert abs(b - want) <= tol, (b, want)
                pa, pb = both(st, codes, coeffs, abi.PAULI_PHYSICAL)
                assert abs(pa - pb) <= 1e-12
    finally:
        modes.restore_defaults()


def test_noalloc_refuses_oversized_sum(hip):
    say("noalloc oversized sum")
    n = 8
    rng = np.random.default_rng(7800)
    psi = bf.random_state(n, rng)
    st = lq.State(n)
    st.set_vector(psi)
    codes, coeffs = random_sum(n, 200000, rng)  # 200k terms: a 4.8 MB term table
    codes, coeffs, pc, pk = _args(codes, coeffs)
    out = ctypes.c_double(12345.0)
    lib = lq.load()
    rc = lib.lq_expect_pauli_sum_noalloc(st.handle, pc, pk, len(coeffs), abi.PAULI_REFERENCE, ctypes.byref(out))
    assert rc == abi.LQ_ERR_INVALID_ARGUMENT
    assert b"static workspace" in lib.lq_last_error()
    assert out.value == 12345.0
    # later calls are unaffected
    small, scoef = random_sum(n, 9, rng)
    a, b = both(st, small, scoef)
    assert abs(a - b) <= 1e-12
    assert abs(b - oracle.pauli_sum(psi, n, small, scoef)) <= 1e-10
