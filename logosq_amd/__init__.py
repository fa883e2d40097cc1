"""logosq_amd: MI355X-native dense state-vector backend with LogosQ's API.

Module layout mirrors the reference crate (src/states, src/gates,
src/circuits, src/optimization, src/algorithms).  All amplitude work runs in
the HIP library liblogosq_hip.so; nothing here computes on the CPU.
"""
from . import abi
from ._lib import (LogosQError, InvalidQubitIndex, InvalidStateDimension, DimensionMismatch,
                   ParameterCountMismatch, CircuitQubitMismatch, DeviceError, load)
from .states import State, C128, C64
from .gates import *  # noqa: F401,F403
from .circuits import Circuit, Operation, Param, set_default_exec_mode, default_exec_mode
from .optimization import (Pauli, PauliTerm, PauliObservable, Observable, Ansatz, HardwareEfficientAnsatz,
                           RealAmplitudesAnsatz, EfficientSU2Ansatz, ParameterizedCircuit, EntanglingGate,
                           EntanglingPattern, ParameterShift, AdjointGradient, FiniteDifference, GradientDescent, Adam, VQE,
                           heisenberg_ring, ry_rz_cz_ring_ansatz, vqe_adam_batch)
from . import algorithms
from . import noise
from .noise import (NoiseModel, NoiseParams, DepolarizingNoise, AmplitudeDampingNoise, PhaseDampingNoise,
                    ThermalRelaxationNoise, CompositeNoise)
from .batch import batch_expectations
from . import jit

__version__ = "0.1.0"
