"""``simulate`` orchestration for instruction-list programs
(src/BraketSimulator.jl:97-149, :264-395).  OpenQASM parsing stays in the
reference's front end; a ``Program`` here is the parsed instruction list."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .results import Expectation, Probability, Sample, Variance, calculate
from .simulators import Instruction, Measure, StateVectorSimulator


@dataclass
class Program:
    instructions: list
    results: list = field(default_factory=list)
    basis_rotation_instructions: list = field(default_factory=list)
    # the OpenQASM text and inputs the instructions were parsed from, echoed
    # into a task result's action (src/BraketSimulator.jl:430-433)
    source: str | None = None
    inputs: dict | None = None

    def qubit_count(self) -> int:
        qs = set()
        for ins in self.instructions + self.basis_rotation_instructions:
            qs.update(ins.target)
        return (max(qs) + 1) if qs else 0


@dataclass
class TaskResult:
    measurements: list
    measured_qubits: list
    result_types: list
    values: list
    shots: int


def index_to_endian_bits(ix: int, n: int) -> list:
    """src/BraketSimulator.jl:88-95."""
    return [(ix >> (n - 1 - k)) & 1 for k in range(n)]


def formatted_measurements(sim, measured_qubits, seed=None) -> np.ndarray:
    """_formatted_measurements (src/BraketSimulator.jl:97-101): shots -> bit
    rows of the measured qubits (qubit 0 = MSB), unpacked on the GPU."""
    shots = sim.samples(seed=seed)
    mq = np.array(measured_qubits, dtype=np.int32)
    out = np.empty((len(shots), len(mq)), dtype=np.int8)
    L.call("bsim_shots_to_bits", sim.handle, len(shots), L.i64ptr(shots), len(mq), L.iptr(mq),
           out.ctypes.data_as(L.C.POINTER(L.C.c_int8)))
    return out


def _measured_qubits(program: Program, n: int) -> list:
    """_get_measured_qubits (src/BraketSimulator.jl:193-200)."""
    mq = []
    for ins in program.instructions:
        if isinstance(ins.operator, Measure):
            for t in ins.target:
                if t not in mq:
                    mq.append(t)
    return mq or list(range(n))


def simulate(sim, program: Program, shots: int = 0, seed=None, evolve_mode: int = 0) -> TaskResult:
    """simulate(simulator, program, shots) (src/BraketSimulator.jl:307-322)."""
    n = program.qubit_count()
    ops = list(program.instructions)
    if shots > 0 and program.basis_rotation_instructions:
        ops += list(program.basis_rotation_instructions)
    sim.reinit(n, shots)
    sim.evolve(ops, mode=evolve_mode)
    mq = _measured_qubits(program, n)
    values = []
    if shots == 0 and program.results:
        values = [calculate(rt, sim) for rt in program.results]
    elif program.results:
        values = [0.0 for _ in program.results]
    meas = formatted_measurements(sim, mq, seed) if shots > 0 else np.empty((0, len(mq)), dtype=np.int8)
    return TaskResult(meas, mq, list(program.results), values, shots)


def simulate_batch(sim, programs, shots: int = 0, seed=None, persistent: bool = True):
    """Batch simulate (src/BraketSimulator.jl:347-395).  A state-vector batch
    with shots whose programs are gate lists of 4..26 qubits runs as one
    persistent launch (bsim_batch_run; shots by inverse CDF, see bsim.batch);
    any other batch runs program by program on ``sim``."""
    from .batch import run_batch
    from .gates import Gate

    def eligible(p):
        n = p.qubit_count()
        ok = 4 <= n <= 26
        for ins in p.instructions + p.basis_rotation_instructions:
            ok = ok and (isinstance(ins.operator, (Gate, Measure)) or ins.operator is None)
        return ok

    if persistent and shots > 0 and isinstance(sim, StateVectorSimulator) and programs and all(map(eligible, programs)):
        circuits, mqs = [], []
        for p in programs:
            n = p.qubit_count()
            ops = [i for i in p.instructions + p.basis_rotation_instructions if not isinstance(i.operator, Measure)]
            circuits.append((n, ops))
            mqs.append(_measured_qubits(p, n))
        seeds = None if seed is None else [seed + i for i in range(len(programs))]
        res = run_batch(circuits, shots, seeds)
        out = []
        for p, (n, _), mq, r in zip(programs, circuits, mqs, res):
            sh = np.asarray(mq, dtype=np.int64)
            meas = ((r.samples[:, None] >> (n - 1 - sh)[None, :]) & 1).astype(np.int8)
            out.append(TaskResult(meas, mq, list(p.results), [0.0 for _ in p.results], shots))
        return out
    return [simulate(sim, p, shots, seed=None if seed is None else seed + i) for i, p in enumerate(programs)]
