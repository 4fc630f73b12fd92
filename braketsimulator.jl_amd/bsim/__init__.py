"""bsim -- MI355X (gfx950) engine for BraketSimulator.jl's evolution hot path.

Host-side mirror of the reference's simulator API over the libbsim C ABI
(include/bsim.h).  Every compute call runs in libbsim's HIP kernels.
"""
from . import gates, noises, observables, wire
from ._lib import BsimError, load
from .gates import *  # noqa: F401,F403 (gate types, as the reference exports them)
from .noises import *  # noqa: F401,F403
from .results import Amplitude, DensityMatrix, Expectation, Probability, Sample, StateVector, Variance, calculate
from .simulate import Program, TaskResult, simulate, simulate_batch
from .simulators import (
    Barrier,
    CapturedEvolution,
    Delay,
    DensityMatrixSimulator,
    Instruction,
    Measure,
    Reset,
    StateVectorSimulator,
    Superoperator,
    diagonalizing_gates,
    lower_instructions,
)

__all__ = [
    "BsimError", "load", "gates", "noises", "observables", "StateVectorSimulator", "DensityMatrixSimulator",
    "Instruction", "Measure", "Barrier", "Reset", "Delay", "Program", "TaskResult", "simulate", "simulate_batch",
    "Expectation", "Variance", "Probability", "Sample", "StateVector", "DensityMatrix", "Amplitude", "calculate",
    "diagonalizing_gates", "lower_instructions", "CapturedEvolution", "Superoperator",
]
