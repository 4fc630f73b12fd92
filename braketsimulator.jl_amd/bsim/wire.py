"""Result wire format: GateModelTaskResult as JSON text, with large result
values handed over through files (the string API the reference exposes to
an out-of-process caller, src/BraketSimulator.jl:397-457).

  - ``lower_observable`` / ``lower_result_type``: the IR forms the result
    types carry (src/observables.jl:37,72,126; src/results.jl:44,77,197-198);
  - ``lower_value``: ResultTypeValue's value lowering (complex -> [re, im],
    src/raw_schema.jl:126-135, complex matrices through complex_matrix_to_ir,
    src/BraketSimulator.jl:48-58);
  - ``task_result``: GateModelTaskResult / TaskMetadata / AdditionalMetadata
    (src/BraketSimulator.jl:103-149, src/raw_schema.jl:231-254);
  - ``mmap_large_result_values``: values above 2^20 bytes are written raw to
    temporary files and emptied in the result (src/BraketSimulator.jl:410-428);
  - ``simulate_json``: simulate(simulator_id, program(s), shots) ->
    (JSON text, file paths, value lengths) (src/BraketSimulator.jl:430-457).

Measurement rows (shots x measured qubits of 0/1) are written as JSON text
by array operations on the bit matrix the device unpacked, so a million-shot
result costs one pass over its bytes rather than a Python loop per shot.
"""
from __future__ import annotations

import json
import os
import tempfile
import uuid

import numpy as np

from . import _lib as L
from . import observables as Obs
from .results import Amplitude, DensityMatrix, Expectation, Probability, Sample, StateVector, Variance, calculate
from .simulate import Program, _measured_qubits, formatted_measurements

_LARGE = 2 ** 20  # bytes (src/BraketSimulator.jl:411)

_DEVICE_ID = {"sv": "braket_sv_v2", "dm": "braket_dm_v2"}


def complex_matrix_to_ir(m) -> list:
    """src/BraketSimulator.jl:48-57: rows of [re, im] pairs."""
    m = np.asarray(m)
    return [[[float(z.real), float(z.imag)] for z in row] for row in m]


def lower_observable(obs) -> list:
    """StructTypes.lower of an observable: ["x"] for a standard observable, a
    complex matrix for a Hermitian one, the factors' lowerings concatenated
    for a tensor product (src/observables.jl:37,72,126)."""
    if isinstance(obs, Obs.TensorProduct):
        out = []
        for f in obs.factors:
            out += lower_observable(f)
        return out
    if isinstance(obs, Obs.HermitianObservable):
        return [complex_matrix_to_ir(obs.matrix())]
    for cls, label in ((Obs.X, "x"), (Obs.Y, "y"), (Obs.Z, "z"), (Obs.I, "i"), (Obs.H, "h")):
        if isinstance(obs, cls):
            return [label]
    raise TypeError(f"no IR form for observable {obs!r}")


def _targets(t):
    return [int(q) for q in t] if t else None


def lower_result_type(rt) -> dict:
    """IR form of a result type (src/results.jl:44,77,197-198); empty targets
    lower to null."""
    for cls, label in ((Expectation, "expectation"), (Variance, "variance"), (Sample, "sample")):
        if isinstance(rt, cls):
            return {"observable": lower_observable(rt.observable), "targets": _targets(rt.targets), "type": label}
    for cls, label in ((Probability, "probability"), (DensityMatrix, "densitymatrix")):
        if isinstance(rt, cls):
            return {"targets": _targets(rt.targets), "type": label}
    if isinstance(rt, Amplitude):
        return {"states": list(rt.states), "type": "amplitude"}
    if isinstance(rt, StateVector):
        return {"type": "statevector"}
    raise TypeError(f"no IR form for result type {rt!r}")


def _lower_complex(z) -> list:
    return [float(np.real(z)), float(np.imag(z))]


def lower_value(v):
    """ResultTypeValue value lowering (src/raw_schema.jl:126-135): floats as
    they are, complex numbers as [re, im], complex matrices through
    complex_matrix_to_ir, dicts entry by entry."""
    if isinstance(v, dict):
        return {str(k): lower_value(x) for k, x in v.items()}
    if isinstance(v, (complex, np.complexfloating)):
        return _lower_complex(v)
    if isinstance(v, (float, int, np.floating, np.integer)):
        return float(v)
    a = np.asarray(v)
    if a.ndim == 2 and np.iscomplexobj(a):
        return complex_matrix_to_ir(a)
    if np.iscomplexobj(a):
        return np.stack([a.real, a.imag], axis=-1).astype(np.float64)
    return a.astype(np.float64)


def measurements_json(bits: np.ndarray) -> str:
    """JSON text of a shots x m matrix of 0/1 (one row per shot), built as
    one byte array: '[' b ',' b ... ']' per row, rows joined by ','."""
    bits = np.asarray(bits, dtype=np.uint8)
    s, m = bits.shape if bits.ndim == 2 else (0, 0)
    if s == 0:
        return "[]"
    w = 2 * m + 2 if m else 3  # row text plus the separating comma
    row = np.empty((s, w), dtype=np.uint8)
    row[:, 0] = ord("[")
    if m:
        row[:, 1:2 * m:2] = bits + ord("0")
        row[:, 2:2 * m:2] = ord(",")
        row[:, 2 * m] = ord("]")
    else:
        row[:, 1] = ord("]")
    row[:, w - 1] = ord(",")
    body = row.reshape(-1)[:-1].tobytes().decode("ascii")
    return "[" + body + "]"


def _action(program: Program) -> dict:
    source = getattr(program, "source", None)
    if source is not None:
        return {"braketSchemaHeader": {"name": "braket.ir.openqasm.program", "version": "1"}, "source": source,
                "inputs": getattr(program, "inputs", None)}
    return {"braketSchemaHeader": {"name": "braket.ir.jaqcd.program", "version": "1"},
            "instructions": [{"type": type(i.operator).__name__, "targets": [int(t) for t in i.target]}
                             for i in program.instructions],
            "results": [lower_result_type(r) for r in program.results],
            "basis_rotation_instructions": [{"type": type(i.operator).__name__, "targets": [int(t) for t in i.target]}
                                            for i in program.basis_rotation_instructions]}


def task_result(sim, program: Program, shots: int, values: list, measurements, measured_qubits) -> dict:
    """GateModelTaskResult (src/BraketSimulator.jl:127-149): sorted measured
    qubits, the formatted measurements (null without shots), result types
    with their lowered values, task metadata (a fresh id, shots, device id)
    and the program as the action."""
    kind = "sv" if sim.kind == L.BSIM_SV else "dm"
    return {
        "braketSchemaHeader": {"name": "braket.task_result.gate_model_task_result", "version": "1"},
        "measurements": measurements,
        "measurementProbabilities": None,
        "resultTypes": [{"type": lower_result_type(rt), "value": lower_value(v)} for rt, v in zip(program.results, values)],
        "measuredQubits": sorted(int(q) for q in measured_qubits),
        "outputs": None,
        "taskMetadata": {
            "braketSchemaHeader": {"name": "braket.task_result.task_metadata", "version": "1"},
            "id": str(uuid.uuid4()), "shots": int(shots), "deviceId": _DEVICE_ID[kind], "deviceParameters": None,
            "createdAt": None, "endedAt": None, "status": None, "failureReason": None,
        },
        "additionalMetadata": {
            "action": _action(program), "dwaveMetadata": None, "ionqMetadata": None, "rigettiMetadata": None,
            "oqcMetadata": None, "xanaduMetadata": None, "queraMetadata": None, "simulatorMetadata": None,
        },
    }


def mmap_large_result_values(result: dict, directory: str | None = None):
    """Values of more than 2^20 bytes go to temporary files as raw
    little-endian float64 data (complex values as re, im pairs), and the
    result keeps an empty list; returns (paths, lengths) with the element
    count of each value, or (None, None) when nothing is large
    (src/BraketSimulator.jl:410-428)."""
    paths, lens = [], []
    for rtv in result["resultTypes"]:
        v = rtv["value"]
        if isinstance(v, np.ndarray) and v.nbytes > _LARGE:
            fd, path = tempfile.mkstemp(prefix="bsim_result_", dir=directory)
            with os.fdopen(fd, "wb") as f:
                np.ascontiguousarray(v, dtype="<f8").tofile(f)
            lens.append(int(v.shape[0]))
            paths.append(path)
            rtv["value"] = []
    if not paths:
        return None, None
    return tuple(paths), tuple(lens)


class _Encoder(json.JSONEncoder):
    def default(self, o):
        if isinstance(o, np.ndarray):
            return o.tolist()
        return super().default(o)


_MEAS = "\u0000bsim-measurements\u0000"


def to_json(result: dict, measurements_text: str | None = None) -> str:
    """JSON text of a task result; ``measurements_text`` (from
    measurements_json) is spliced in for the measurements field."""
    if measurements_text is None:
        return json.dumps(result, cls=_Encoder)
    r = dict(result)
    r["measurements"] = _MEAS
    text = json.dumps(r, cls=_Encoder)
    return text.replace(json.dumps(_MEAS), measurements_text, 1)


def _run_one(sim, program: Program, shots: int, seed, directory):
    n = program.qubit_count()
    ops = list(program.instructions)
    if shots > 0 and program.basis_rotation_instructions:
        ops += list(program.basis_rotation_instructions)
    sim.reinit(n, shots)
    sim.evolve(ops)
    mq = sorted(_measured_qubits(program, n))
    if shots == 0:
        values = [calculate(rt, sim) for rt in program.results]
    else:
        values = [0.0 for _ in program.results]  # the caller computes these from the shots
    meas_text = measurements_json(formatted_measurements(sim, mq, seed)) if shots > 0 else None
    res = task_result(sim, program, shots, values, None, mq)
    paths, lens = mmap_large_result_values(res, directory)
    return to_json(res, meas_text), paths, lens


def create_sim(simulator_id: str, shots: int = 0):
    """src/BraketSimulator.jl:402-408."""
    from .simulators import DensityMatrixSimulator, StateVectorSimulator

    if simulator_id == "braket_sv_v2":
        return StateVectorSimulator(0, shots)
    if simulator_id == "braket_dm_v2":
        return DensityMatrixSimulator(0, shots)
    raise ValueError(f"unknown simulator id {simulator_id!r}")


def simulate_json(simulator_id: str, programs, shots: int = 0, seed=None, directory: str | None = None):
    """simulate(simulator_id, task_spec(s), inputs, shots) for the string API
    (src/BraketSimulator.jl:430-457).  One Program: (JSON text, paths,
    lengths).  A list: (JSON array text, JSON text of the per-result
    [paths, lengths] pairs)."""
    sim = create_sim(simulator_id, shots)
    if isinstance(programs, Program):
        return _run_one(sim, programs, shots, seed, directory)
    texts, pl = [], []
    for i, p in enumerate(programs):
        t, paths, lens = _run_one(sim, p, shots, None if seed is None else seed + i, directory)
        texts.append(t)
        pl.append([list(paths) if paths else None, list(lens) if lens else None])
    return "[" + ",".join(texts) + "]", json.dumps(pl)


def load_result(text: str, paths=None, lens=None) -> dict:
    """Parse a task result and read its file-backed values back in (the
    consumer side of mmap_large_result_values: the files hold float64 data,
    complex values as pairs)."""
    res = json.loads(text)
    if paths:
        big = [r for r in res["resultTypes"] if r["value"] == []]
        for r, path, n in zip(big, paths, lens):
            a = np.fromfile(path, dtype="<f8")
            r["value"] = a.reshape(n, -1) if a.size != n else a
    return res
