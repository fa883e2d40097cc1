"""The run-time specialised expectation kernels on the MI355X against the
oracle (PauliObservable::expectation, src/optimization/observable.rs:95-147):
XYZ rings and local sums with Y, XZ and 3-local strings, diagonal terms and
strings too wide for a tile, on states of one tile to many tiles per block;
<= 1e-10 complex128, <= 1e-5 complex64 (f64 sums), both Y conventions, and
equal to the generic kernels' result to rounding.  The kernels are compiled
at first sight (LQ_JIT_ALWAYS) and used at every size (tiles on)."""
import ctypes

import numpy as np
import pytest

import oracle
import logosq_amd as lq
from logosq_amd import abi, algorithms, jit
from tests import bruteforce as bf
from tests import modes
from tests.test_expect_jit_emul import local_sum

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def expect(st, codes, coeffs, conv=abi.PAULI_REFERENCE):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
    e = ctypes.c_double()
    lq._lib.call("lq_expect_pauli_sum", st.handle, codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                 coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(coeffs), conv, ctypes.byref(e))
    return e.value


def layouts(n, dtype, codes, coeffs):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
    lay, pair, diag = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lq._lib.call("lq_expect_plan", n, dtype, codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                 coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(coeffs), abi.PAULI_REFERENCE,
                 ctypes.byref(lay), ctypes.byref(pair), ctypes.byref(diag))
    return lay.value


@pytest.mark.parametrize("n", [10, 13, 16, 21])
@pytest.mark.parametrize("kind", ["ring", "local"])
def test_specialised_expectation_matches_oracle(hip, n, kind):
    say("specialised expectation %s n=%d" % (kind, n))
    rng = np.random.default_rng(9300 + n + (100 if kind == "local" else 0))
    if kind == "ring":
        codes, coeffs = algorithms.heisenberg_codes(n, *rng.uniform(-1, 1, 3))
    else:
        codes, coeffs = local_sum(n, rng)
    psi = bf.random_state(n, rng)
    want = oracle.pauli_sum(psi, n, codes, coeffs)
    want_phys = oracle.pauli_sum(psi, n, codes, coeffs * ((-1.0) ** (codes == 2).sum(axis=1)))
    sts = {}
    for dtype in (lq.C128, lq.C64):
        sts[dtype] = lq.State(n, dtype)
        sts[dtype].set_vector(psi)
    try:
        jit.set_jit(jit.OFF)
        lq._lib.call("lq_set_expect_tiles", abi.EXPECT_TILES_ON)
        generic = expect(sts[lq.C128], codes, coeffs)
        jit.set_jit(jit.ALWAYS)
        before = jit.jit_stats()["failed"]
        for dtype, tol in ((lq.C128, 1e-10), (lq.C64, 1e-5)):
            assert layouts(n, dtype, codes, coeffs) >= 1
            got = expect(sts[dtype], codes, coeffs)
            phys = expect(sts[dtype], codes, coeffs, abi.PAULI_PHYSICAL)
            say("  dtype %d: |got - oracle| %.3e" % (dtype, abs(got - want)))
            assert abs(got - want) <= tol
            assert abs(phys - want_phys) <= tol
            if dtype == lq.C128:
                assert abs(got - generic) <= 1e-12
            again = expect(sts[dtype], codes, coeffs)  # the cached kernels, coefficients re-uploaded
            assert again == got
        assert jit.jit_stats()["failed"] == before
    finally:
        modes.restore_defaults()


def test_specialised_expectation_new_coefficients(hip):
    """One compiled structure serves other coefficients: the same ring with
    new couplings (the kernels read them from the table) still matches."""
    say("specialised expectation, new couplings")
    n = 16
    rng = np.random.default_rng(9400)
    psi = bf.random_state(n, rng)
    st = lq.State(n)
    st.set_vector(psi)
    try:
        jit.set_jit(jit.ALWAYS)
        lq._lib.call("lq_set_expect_tiles", abi.EXPECT_TILES_ON)
        for _ in range(3):
            codes, coeffs = algorithms.heisenberg_codes(n, *rng.uniform(-1, 1, 3))
            assert abs(expect(st, codes, coeffs) - oracle.pauli_sum(psi, n, codes, coeffs)) <= 1e-10
    finally:
        modes.restore_defaults()
