"""The run-time specialised tile-pass kernels compile (hipRTC, no device
needed) from the embedded headers for the plans the fused engine makes, and
the mode switch rejects unknown modes."""
import pytest

import logosq_amd as lq
from logosq_amd import abi, algorithms, jit
from tests import modes
from tests.test_planner import random_circuit


def test_jit_mode_rejects_unknown():
    with pytest.raises(lq.LogosQError):
        jit.set_jit(7)
    modes.restore_defaults()


@pytest.mark.parametrize("dtype", [abi.LQ_C128, abi.LQ_C64])
def test_jit_compiles_brickwork_plan(dtype):
    d = algorithms.brickwork_circuit(14, 3, 1).descriptors()
    before = jit.jit_stats()
    k = jit.compile_plan(14, d, dtype)
    after = jit.jit_stats()
    assert k >= 1
    assert after["failed"] == before["failed"]


def test_jit_compiles_every_op_kind():
    """Random circuits hold every gate kind (pairs of each matrix class,
    diagonals with selectors inside and outside the subcube, SWAP, 4x4)."""
    import numpy as np
    rng = np.random.default_rng(3)
    gates = random_circuit(13, 120, rng, with_dense=False)
    assert jit.compile_plan(13, gates, abi.LQ_C128) >= 1


@pytest.mark.parametrize("dtype", [abi.LQ_C128, abi.LQ_C64])
def test_jit_compiles_sign_ops(dtype):
    """Z and CZ with controls inside and outside the group subcubes (sign
    flips at compile time and by per-thread flags) compile."""
    import numpy as np
    from tests.test_gpu_jit import sign_heavy_circuit
    gates = sign_heavy_circuit(14, 3, np.random.default_rng(5))
    before = jit.jit_stats()
    assert jit.compile_plan(14, gates, dtype) >= 1
    assert jit.jit_stats()["failed"] == before["failed"]
