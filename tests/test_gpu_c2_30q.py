"""The brickwork workload at its full size: 30 qubits, depth 20, seed 2
(890 gates: 600 random RX/RY/RZ, 290 CNOTs), from |0...0>, complex128 and
complex64.  The oracle cannot run 2^30 amplitudes in a test's time, so the
checks are the size-independent ones:

* per-gate execution (one kernel per gate, Gate::apply) against the fused
  tile passes on their run-time specialised kernels, at the default tile
  geometry and at the 13-bit tiles the autotuner picks: max |delta| <= 1e-12;
* the circuit followed by its inverse returns |0...0>: |<0|psi> - 1| <= 1e-12
  (complex128, per-gate and fused);
* the XYZ-ring energy (90 terms) from LDS tiles against one read per flip
  group: <= 1e-10;
* complex64 (per-gate and fused) against the complex128 state, with the
  error bounds written below and the measured values printed.

Both paths are checked against the oracle at 22 qubits in
tests/test_gpu_fused.py; the reference's execute loop is
src/circuits/mod.rs:352-375.  Host memory: two 16 GiB buffers."""
import ctypes
import math

import numpy as np
import pytest

import logosq_amd as lq
from logosq_amd import abi, jit
from tests import modes

pytestmark = pytest.mark.gpu

N = 30
DEPTH = 20
SEED = 2
J = (0.37, -0.81, 0.55)

# complex64 against complex128 (f32 storage, 890 gates): bounds on the
# largest amplitude error, the relative 2-norm error and the round trip.
C64_MAX_ABS = 1e-8
C64_REL_L2 = 1e-5
C64_ROUND_TRIP = 1e-5
C64_ENERGY = 1e-5


def say(*a):
    print(*a, flush=True)


def inverse(gates):
    """The inverse circuit: reversed order, rotation angles negated (CNOT is
    its own inverse)."""
    out = []
    for g in reversed(gates):
        h = abi.LqGate()
        ctypes.pointer(h)[0] = g
        if h.kind in (abi.RX, abi.RY, abi.RZ):
            h.theta[0] = -h.theta[0]
        else:
            assert h.kind == abi.CNOT
        out.append(h)
    return out


def chunks(n, size=1 << 26):
    for s in range(0, n, size):
        yield slice(s, min(n, s + size))


def max_abs_diff(a, b):
    return max(float(np.max(np.abs(a[s] - b[s]))) for s in chunks(len(a)))


def l2_diff(a, b):
    return math.sqrt(sum(float(np.sum(np.abs(a[s] - b[s]) ** 2)) for s in chunks(len(a))))


class C2:
    """States and host buffers shared by the tests of this module, built on
    first use so that each test also runs on its own."""

    def __init__(self):
        circ = lq.algorithms.brickwork_circuit(N, DEPTH, SEED)
        self.gates = circ.descriptors()
        assert len(self.gates) == 890
        self.inv = inverse(self.gates)
        self.arr = abi.gate_array(self.gates)
        self.inv_arr = abi.gate_array(self.inv)
        self.per_gate = None   # complex128 state after the per-gate run
        self.host = None       # its amplitudes on the host
        self.buf = None        # second host buffer

    def run(self, st, arr, count, mode):
        lq._lib.call("lq_execute", st.handle, arr, count, mode)
        lq._lib.call("lq_synchronize")

    def per_gate_state(self):
        if self.per_gate is None:
            say("30q c128 per-gate forward")
            st = lq.State(N)
            self.run(st, self.arr, len(self.gates), abi.EXEC_PER_GATE)
            self.per_gate = st
            self.host = np.empty(1 << N, dtype=np.complex128)
            self.download(st, self.host)
        return self.per_gate

    def reference_host(self):
        self.per_gate_state()
        return self.host

    def scratch(self):
        if self.buf is None:
            self.buf = np.empty(1 << N, dtype=np.complex128)
        return self.buf

    @staticmethod
    def download(st, out):
        lq._lib.call("lq_state_download", st.handle, out.view(np.float64).ctypes.data_as(
            ctypes.POINTER(ctypes.c_double)), len(out))


@pytest.fixture(scope="module")
def c2(hip):
    lq.load()
    obj = C2()
    jit.set_jit(jit.ALWAYS)
    yield obj
    lq._lib.call("lq_set_fused_config", 0, 0, 0)
    modes.restore_defaults()


def test_c2_30q_c128_fused_vs_per_gate(c2):
    ref = c2.reference_host()
    before = jit.jit_stats()
    st = lq.State(N)
    buf = c2.scratch()
    for tile in (0, 13):  # default geometry (12-bit tiles), then the autotuner's 13
        say("30q c128 fused, tile bits %s" % (tile or "default"))
        lq._lib.call("lq_set_fused_config", tile, 0, 0)
        lq._lib.call("lq_state_set_basis", st.handle, 0)
        c2.run(st, c2.arr, len(c2.gates), abi.EXEC_FUSED)
        c2.download(st, buf)
        err = max_abs_diff(buf, ref)
        say("  max |fused - per-gate| = %.3e" % err)
        assert err <= 1e-12, (tile, err)
    lq._lib.call("lq_set_fused_config", 0, 0, 0)
    after = jit.jit_stats()
    # the passes ran on their specialised kernels (none fell back)
    assert after["launches"] > before["launches"]
    assert after["failed"] == before["failed"]
    nrm = sum(float(np.sum(np.abs(ref[s]) ** 2)) for s in chunks(len(ref)))
    say("  norm^2 = %.15f" % nrm)
    assert abs(nrm - 1.0) <= 1e-12


def test_c2_30q_xyz_energy_tiles_vs_groups(c2):
    st = c2.per_gate_state()
    obs = lq.heisenberg_ring(N, *J)
    codes, coeffs = obs.codes()
    tl, pg, dg = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    out = {}
    try:
        for mode in (abi.EXPECT_TILES_ON, abi.EXPECT_TILES_OFF):
            lq._lib.call("lq_set_expect_tiles", mode)
            lq._lib.call("lq_expect_plan", N, lq.C128, codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                         coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(coeffs),
                         abi.PAULI_REFERENCE, ctypes.byref(tl), ctypes.byref(pg), ctypes.byref(dg))
            out[mode] = obs.expectation(st)
            say("30q XYZ energy, tiles %s: %.17g (%d layouts, %d pair groups, %d diagonal groups)"
                % ("on" if mode == abi.EXPECT_TILES_ON else "off", out[mode], tl.value, pg.value, dg.value))
            if mode == abi.EXPECT_TILES_ON:
                assert tl.value > 0
            else:
                assert tl.value == 0
    finally:
        modes.restore_defaults()
    assert abs(out[abi.EXPECT_TILES_ON] - out[abi.EXPECT_TILES_OFF]) <= 1e-10


def test_c2_30q_c64_against_c128(c2):
    ref = c2.reference_host()
    buf = c2.scratch()
    obs = lq.heisenberg_ring(N, *J)
    e128 = obs.expectation(c2.per_gate_state())
    st = lq.State(N, lq.C64)
    for mode in (abi.EXEC_PER_GATE, abi.EXEC_FUSED):
        name = "per-gate" if mode == abi.EXEC_PER_GATE else "fused"
        say("30q c64 %s forward" % name)
        lq._lib.call("lq_state_set_basis", st.handle, 0)
        c2.run(st, c2.arr, len(c2.gates), mode)
        c2.download(st, buf)
        err = max_abs_diff(buf, ref)
        rel = l2_diff(buf, ref)  # the reference state has norm 1
        e64 = obs.expectation(st)
        say("  c64 %s vs c128: max |delta| %.3e, relative 2-norm %.3e, energy delta %.3e"
            % (name, err, rel, abs(e64 - e128)))
        assert err <= C64_MAX_ABS
        assert rel <= C64_REL_L2
        assert abs(e64 - e128) <= C64_ENERGY


def test_c2_30q_round_trip(c2):
    """Forward then inverse returns |0...0>, per-gate and fused (c128), and
    fused in complex64."""
    for dtype, mode, tol in ((lq.C128, abi.EXEC_FUSED, 1e-12), (lq.C128, abi.EXEC_PER_GATE, 1e-12),
                             (lq.C64, abi.EXEC_FUSED, C64_ROUND_TRIP)):
        name = "%s %s" % ("c128" if dtype == lq.C128 else "c64", "fused" if mode == abi.EXEC_FUSED else "per-gate")
        say("30q round trip %s" % name)
        if dtype == lq.C128 and mode == abi.EXEC_PER_GATE and c2.per_gate is not None:
            st = c2.per_gate  # the forward state is already there; the host copy stays
            c2.per_gate = None
        else:
            st = lq.State(N, dtype)
            c2.run(st, c2.arr, len(c2.gates), mode)
        c2.run(st, c2.inv_arr, len(c2.inv), mode)
        zero = lq.State(N, dtype)
        amp0 = st.inner_product(zero).conjugate()  # <0|psi>
        nrm = st.norm_sqr()
        say("  |<0|psi> - 1| = %.3e, |norm^2 - 1| = %.3e" % (abs(amp0 - 1.0), abs(nrm - 1.0)))
        assert abs(amp0 - 1.0) <= tol, name
        del zero, st
