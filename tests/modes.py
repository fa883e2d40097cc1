"""The run-time paths' session modes: the library defaults, or with
LQ_TEST_FORCE_PATHS=1 the specialised pass kernels at first sight and the
LDS-tile expectation at every size (tests that switch a mode restore these)."""
import os

import logosq_amd as lq
from logosq_amd import abi, jit


def forced():
    return os.environ.get("LQ_TEST_FORCE_PATHS") == "1"


def restore_defaults():
    lq.load()
    jit.set_jit(jit.ALWAYS if forced() else jit.REPEAT)
    lq._lib.call("lq_set_expect_tiles", abi.EXPECT_TILES_ON if forced() else abi.EXPECT_TILES_AUTO)
    lq._lib.call("lq_set_gate_unit_order", -1)
