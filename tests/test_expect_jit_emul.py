"""The run-time specialised expectation kernels, run on the CPU: the library
plans a Pauli sum (lq_expect_jit_describe gives each layout's kernel text and
coefficient table), each kernel text is compiled as host code against the
stand-in header in tests/emul (one thread per lane, __syncthreads as a
barrier, AddressSanitizer on every access) and run over a random state; the
layouts' sum plus the groups left to the pair kernel must equal the oracle's
PauliObservable::expectation (src/optimization/observable.rs:95-147):
<= 1e-10 complex128, <= 1e-5 complex64 (f32 state, f64 sums).  No device is
used; the GPU tests run the same kernels on the MI355X."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
import logosq_amd as lq
from logosq_amd import abi, algorithms, jit
from tests import bruteforce as bf
from tests import modes

HERE = os.path.dirname(os.path.abspath(__file__))
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(autouse=True)
def specialised_modes():
    lq.load()
    jit.set_jit(jit.ALWAYS)
    lq._lib.call("lq_set_expect_tiles", abi.EXPECT_TILES_ON)
    yield
    modes.restore_defaults()


def describe(n, dtype, codes, coeffs):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
    need = ctypes.c_size_t()
    args = (n, dtype, codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
            coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(coeffs), abi.PAULI_REFERENCE)
    lq._lib.call("lq_expect_jit_describe", *args, None, 0, ctypes.byref(need))
    buf = ctypes.create_string_buffer(need.value)
    lq._lib.call("lq_expect_jit_describe", *args, buf, need.value, None)
    return json.loads(buf.value.decode())


def flip_groups(n, codes):
    """Term indices per flip mask, groups in first-appearance order (the
    library's grouping)."""
    order, members = [], {}
    for t, row in enumerate(codes):
        x = 0
        for q in range(n):
            if row[q] in (1, 2):
                x |= 1 << (n - 1 - q)
        if x not in members:
            members[x] = []
            order.append(x)
        members[x].append(t)
    return order, members


def run_layouts(tmp_path, plan, psi, dtype):
    # the compiler and the emulator carry their own sanitizer runtime: no
    # preloaded one (a sanitized library run preloads it into this process)
    env = {k: v for k, v in os.environ.items() if k != "LD_PRELOAD"}
    state = tmp_path / "state.bin"
    np.ascontiguousarray(psi, dtype=np.complex128).view(np.float64).tofile(state)
    total = 0.0
    for k, lay in enumerate(plan["layouts"]):
        src = tmp_path / ("layout%d.hip" % k)
        src.write_text(lay["src"])
        coef = tmp_path / ("coef%d.bin" % k)
        np.array(lay["coef"] or [0.0], dtype=np.float64).tofile(coef)
        exe = tmp_path / ("layout%d" % k)
        cmd = [CLANG, "-std=c++20", "-O1", "-g", "-fsanitize=address", "-pthread", "-I", os.path.join(HERE, "emul"),
               '-DLQ_KERNEL_FILE="%s"' % src, "-DLQ_V=%s" % ("double2" if dtype == lq.C128 else "float2"),
               os.path.join(HERE, "emul", "run_layout.cpp"), "-o", str(exe)]
        r = subprocess.run(cmd, capture_output=True, text=True, env=env)
        assert r.returncode == 0, r.stderr[-3000:]
        r = subprocess.run([str(exe), str(plan["n"]), str(dtype), str(plan["blocks"]), str(lay["threads"]), str(state),
                            str(coef)], capture_output=True, text=True, timeout=300, env=env)
        assert r.returncode == 0, r.stderr[-3000:]
        total += float(r.stdout)
    return total


def rest_part(n, codes, coeffs, psi, plan):
    order, members = flip_groups(n, codes)
    terms = [t for g in plan["rest"] for t in members[order[g]]]
    if not plan["diag_in"] and 0 in members:
        terms += members[0]
    if not terms:
        return 0.0
    return oracle.pauli_sum(psi, n, codes[terms], coeffs[terms])


def local_sum(n, rng):
    """Ring bonds of every Pauli pair, 3-local strings, a few diagonal
    terms (Z, ZZ far apart) and two long random strings."""
    rows, cs = [], []
    for q in range(n):
        for a, b in ((1, 1), (2, 2), (3, 3), (1, 3), (2, 1)):
            r = np.zeros(n, np.uint8)
            r[q], r[(q + 1) % n] = a, b
            rows.append(r)
        r = np.zeros(n, np.uint8)
        r[q], r[(q + 2) % n], r[(q + 5) % n] = 2, 3, 1
        rows.append(r)
    for q in range(0, n, 3):
        r = np.zeros(n, np.uint8)
        r[q] = 3
        r[(q + n // 2) % n] = 3
        rows.append(r)
    for _ in range(2):
        rows.append(rng.integers(0, 4, n).astype(np.uint8))
    rows = np.array(rows)
    return rows, rng.normal(size=len(rows))


CASES = [
    ("ring", 13, lq.C128), ("ring", 15, lq.C128), ("ring", 14, lq.C64), ("ring", 16, lq.C64),
    ("local", 13, lq.C128), ("local", 16, lq.C128), ("local", 15, lq.C64),
]


@pytest.mark.parametrize("kind,n,dtype", CASES)
def test_specialised_layouts_match_oracle(tmp_path, kind, n, dtype):
    rng = np.random.default_rng(9100 + 10 * n + dtype)
    if kind == "ring":
        codes, coeffs = algorithms.heisenberg_codes(n, *rng.uniform(-1, 1, 3))
    else:
        codes, coeffs = local_sum(n, rng)
    plan = describe(n, dtype, codes, coeffs)
    assert plan.get("layouts"), plan
    psi = bf.random_state(n, rng)
    got = run_layouts(tmp_path, plan, psi, dtype) + rest_part(n, codes, coeffs, psi, plan)
    want = oracle.pauli_sum(psi, n, codes, coeffs)
    tol = 1e-10 if dtype == lq.C128 else 1e-5
    assert abs(got - want) <= tol, (got, want, [(l["L"], l["windows"], l["pair_groups"]) for l in plan["layouts"]])


def test_xyz_ring_30_qubits_reads():
    """The 30-qubit XYZ ring: three layouts for complex128 (the second keeps
    only 128-byte runs of low bits), four for complex64 (runs of at least 16
    amplitudes), the diagonal group inside the first, nothing left to the
    pair kernel."""
    codes, coeffs = algorithms.heisenberg_codes(30, 0.4, -0.7, 0.2)
    for dtype, reads in ((lq.C128, 3), (lq.C64, 4)):
        plan = describe(30, dtype, codes, coeffs)
        assert len(plan["layouts"]) == reads and plan["diag_in"] and plan["rest"] == [], \
            [(l["L"], l["windows"]) for l in plan["layouts"]]


@pytest.mark.parametrize("dtype", [lq.C128, lq.C64])
def test_specialised_kernels_compile_with_hiprtc(dtype):
    """lq_expect_jit_compile builds every layout kernel with hipRTC (the
    library's compile path; no device), for the 30-qubit ring and for a
    local sum with pair groups and thread / tile signs."""
    rng = np.random.default_rng(77)
    for n, (codes, coeffs) in ((30, algorithms.heisenberg_codes(30, 0.7, -0.4, 0.9)), (22, local_sum(22, rng))):
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        before = jit.jit_stats()["failed"]
        k = ctypes.c_int()
        lq._lib.call("lq_expect_jit_compile", n, dtype, codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                     coeffs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(coeffs), abi.PAULI_REFERENCE,
                     ctypes.byref(k))
        assert k.value == len(describe(n, dtype, codes, coeffs)["layouts"]) >= 1
        assert jit.jit_stats()["failed"] == before
