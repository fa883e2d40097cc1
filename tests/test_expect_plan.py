"""The expectation planner (host code, no GPU): which flip groups are read
from LDS tile layouts and how many reads the others take (orbit items: up
to four linearly independent flip masks per read), in each mode.  Generic tile kernels
(run-time specialisation off): a layout holds the 5 lowest bits and 6
(complex128) / 7 (complex64) more, so a ring's bonds pack about 6 to a
layout.  Specialised kernels (the default): 128 KiB tiles whose low bits may
shrink to one 128-byte line, the diagonal group inside the first layout, so
the 30-qubit ring takes three reads (complex64: four).  Strings whose flip bits span more than
a tile are read alone in both."""
import ctypes

import numpy as np
import pytest

import logosq_amd as lq
from logosq_amd import abi, algorithms, jit
from tests import modes


def plan(n, codes, coeffs, dtype=abi.LQ_C128):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
    lay, pair, diag = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lq._lib.call("lq_expect_plan", n, dtype, codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                 coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(coeffs), 0, ctypes.byref(lay),
                 ctypes.byref(pair), ctypes.byref(diag))
    return lay.value, pair.value, diag.value


@pytest.fixture
def tiles_mode():
    lq.load()
    lq._lib.call("lq_set_expect_tiles", abi.EXPECT_TILES_AUTO)
    jit.set_jit(jit.OFF)
    yield lambda m: lq._lib.call("lq_set_expect_tiles", m)
    modes.restore_defaults()


def test_ring_30_specialised_three_reads(tiles_mode):
    jit.set_jit(jit.REPEAT)
    codes, coeffs = algorithms.heisenberg_codes(30, 0.7, -0.4, 0.9)
    assert plan(30, codes, coeffs) == (3, 0, 0)
    assert plan(30, codes, coeffs, abi.LQ_C64) == (4, 0, 0)
    assert plan(20, codes[:0], coeffs[:0]) == (0, 0, 0)
    small = algorithms.heisenberg_codes(20, 0.7, -0.4, 0.9)
    assert plan(20, *small) == (0, 5, 1)  # 16 MiB: no tiles; 20 bonds, four per read
    tiles_mode(abi.EXPECT_TILES_OFF)
    assert plan(30, codes, coeffs) == (0, 8, 1)  # 30 bonds, four per read


def test_ring_30_auto_uses_tiles(tiles_mode):
    codes, coeffs = algorithms.heisenberg_codes(30, 0.7, -0.4, 0.9)
    lay, pair, diag = plan(30, codes, coeffs)
    assert (pair, diag) == (0, 1)
    assert 4 <= lay <= 6  # 30 bonds, about 6 per layout
    lay64, pair64, _ = plan(30, codes, coeffs, abi.LQ_C64)
    assert pair64 == 0 and lay64 <= lay


def test_small_state_auto_orbit_reads(tiles_mode):
    codes, coeffs = algorithms.heisenberg_codes(20, 0.7, -0.4, 0.9)
    assert plan(20, codes, coeffs) == (0, 5, 1)  # 16 MiB: no tiles; 20 bonds, four per read


def test_modes_and_wide_strings(tiles_mode):
    n = 14
    codes, coeffs = algorithms.heisenberg_codes(n, 0.7, -0.4, 0.9)
    wide = np.zeros((1, n), np.uint8)
    wide[0, :] = 1  # X on every qubit: flip bits span all 14 bits, fit no 11-bit tile
    codes = np.vstack([codes, wide])
    coeffs = np.append(coeffs, 0.5)
    tiles_mode(abi.EXPECT_TILES_ON)
    lay, pair, diag = plan(n, codes, coeffs)
    assert lay >= 1 and pair == 1 and diag == 1
    tiles_mode(abi.EXPECT_TILES_OFF)
    assert plan(n, codes, coeffs) == (0, 4, 1)  # 14 bonds and the wide string, four per read
    tiles_mode(abi.EXPECT_TILES_ON)
    assert plan(7, *algorithms.heisenberg_codes(7, 1, 1, 1)) == (0, 2, 1)  # under 8 qubits: no tiles
    jit.set_jit(jit.REPEAT)  # specialised: the diagonal group joins the first layout
    lay, pair, diag = plan(n, codes, coeffs)
    assert lay >= 1 and pair == 1 and diag == 0


def test_mode_rejects_unknown(tiles_mode):
    with pytest.raises(lq.LogosQError):
        tiles_mode(5)
