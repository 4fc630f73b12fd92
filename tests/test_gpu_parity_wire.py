"""The string API end to end on the GPU: simulate_json for the programs of
the reference's Python-extension tests (test/test_python_ext.jl:66-116,
instruction lists in place of the OpenQASM text they are parsed from),
batches, shots, and a probability vector large enough (2 MiB) to travel
through a file."""
import json
import os
import sys

import numpy as np
import pytest

import oracle as O
from bsim import gates as G
from bsim import observables as Obs
from bsim import wire
from bsim.results import Amplitude, DensityMatrix, Expectation, Probability, StateVector
from bsim.simulate import Program
from bsim.simulators import Instruction, Measure

pytestmark = pytest.mark.gpu

R2 = 2 ** -0.5


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bell(second, results, source):
    return Program([Instruction(G.H(), [0]), Instruction(second, [0, 1])], results, source=source)


def test_statevector_amplitude_expectation_results():
    src = "h $0;\ncnot $0, $1;\n#pragma braket result state_vector\n"
    _say("wire: state vector")
    text, paths, lens = wire.simulate_json("braket_sv_v2", _bell(G.CNot(), [StateVector()], src), 0)
    res = json.loads(text)
    assert paths is None and lens is None
    assert res["additionalMetadata"]["action"]["source"] == src
    assert res["braketSchemaHeader"]["name"] == "braket.task_result.gate_model_task_result"
    assert res["taskMetadata"]["deviceId"] == "braket_sv_v2" and res["taskMetadata"]["shots"] == 0
    assert len(res["resultTypes"]) == 1 and res["resultTypes"][0]["type"] == {"type": "statevector"}
    np.testing.assert_allclose(np.array(res["resultTypes"][0]["value"]), [[R2, 0], [0, 0], [0, 0], [R2, 0]],
                               atol=1e-15)
    assert res["measurements"] is None and res["measuredQubits"] == [0, 1]

    _say("wire: amplitude")
    res = json.loads(wire.simulate_json("braket_sv_v2", _bell(G.CY(), [Amplitude(["00", "01", "10", "11"])], "a"),
                                        0)[0])
    rt = res["resultTypes"][0]
    assert rt["type"] == {"states": ["00", "01", "10", "11"], "type": "amplitude"}
    assert set(rt["value"]) == {"00", "01", "10", "11"}
    np.testing.assert_allclose(rt["value"]["00"], [R2, 0.0], atol=1e-15)
    np.testing.assert_allclose(rt["value"]["01"], [0.0, 0.0], atol=1e-15)
    np.testing.assert_allclose(rt["value"]["10"], [0.0, 0.0], atol=1e-15)
    np.testing.assert_allclose(rt["value"]["11"], [0.0, R2], atol=1e-15)

    _say("wire: expectation")
    res = json.loads(wire.simulate_json("braket_sv_v2", _bell(G.CY(), [Expectation(Obs.Z(), [0])], "e"), 0)[0])
    rt = res["resultTypes"][0]
    assert rt["type"] == {"observable": ["z"], "targets": [0], "type": "expectation"}
    assert abs(rt["value"]) < 1e-14


def test_density_matrix_result():
    _say("wire: density matrix")
    res = json.loads(wire.simulate_json("braket_dm_v2", _bell(G.CY(), [DensityMatrix()], "d"), 0)[0])
    assert res["taskMetadata"]["deviceId"] == "braket_dm_v2"
    rt = res["resultTypes"][0]
    assert rt["type"] == {"targets": None, "type": "densitymatrix"}
    m = np.array(rt["value"])
    rho = m[..., 0] + 1j * m[..., 1]
    psi = np.array([R2, 0, 0, 1j * R2])
    # the stored matrix, rho^T, as the reference returns it (src/dm_simulator.jl:136-138)
    np.testing.assert_allclose(rho, np.outer(psi, psi.conj()).T, atol=1e-15)


def test_batch_of_programs():
    _say("wire: batch")
    progs = [_bell(G.CNot(), [Probability([0]), Probability()], "p1"),
             _bell(G.CY(), [Expectation(Obs.TensorProduct([Obs.Z(), Obs.Z()]), [0, 1])], "p2")]
    text, pl = wire.simulate_json("braket_sv_v2", progs, 0)
    res = json.loads(text)
    assert [r["additionalMetadata"]["action"]["source"] for r in res] == ["p1", "p2"]
    np.testing.assert_allclose(res[0]["resultTypes"][0]["value"], [0.5, 0.5], atol=1e-15)
    np.testing.assert_allclose(res[0]["resultTypes"][1]["value"], [0.5, 0, 0, 0.5], atol=1e-15)
    assert abs(res[1]["resultTypes"][0]["value"] - 1.0) < 1e-14
    assert json.loads(pl) == [[None, None], [None, None]]


def test_shots_measurements():
    _say("wire: shots")
    n = 5
    ins = [Instruction(G.H(), [q]) for q in range(n)] + [Instruction(G.CNot(), [q, q + 1]) for q in range(n - 1)]
    ins += [Instruction(Measure(), [3]), Instruction(Measure(), [1])]
    text, _, _ = wire.simulate_json("braket_sv_v2", Program(ins, [Probability([1])]), 2000, seed=11)
    res = json.loads(text)
    assert res["measuredQubits"] == [1, 3]
    meas = np.array(res["measurements"])
    assert meas.shape == (2000, 2) and set(np.unique(meas)) <= {0, 1}
    # the same seed through the simulator object gives the same rows
    from bsim.simulate import formatted_measurements
    from bsim.simulators import StateVectorSimulator

    sim = StateVectorSimulator(n, 2000)
    sim.evolve([i for i in ins if not isinstance(i.operator, Measure)])
    np.testing.assert_array_equal(meas, formatted_measurements(sim, [1, 3], 11))
    assert res["resultTypes"][0]["value"] == 0.0  # shots > 0: computed by the caller from the shots


def test_large_probability_through_file(tmp_path):
    n = 18  # 2^18 doubles = 2 MiB > 2^20 bytes
    _say("wire: large result n=18")
    rng = np.random.default_rng(2)
    ins = [Instruction(G.Ry(float(rng.uniform(0, 3))), [q]) for q in range(n)]
    ins += [Instruction(G.CNot(), [q, q + 1]) for q in range(n - 1)]
    text, paths, lens = wire.simulate_json("braket_sv_v2", Program(ins, [Probability(), Expectation(Obs.Z(), [2])]),
                                           0, directory=str(tmp_path))
    assert lens == (2 ** n,) and len(paths) == 1
    res = wire.load_result(text, paths, lens)
    p = res["resultTypes"][0]["value"]
    np.testing.assert_allclose(p, O.sv_probabilities(O.evolve_sv(n, ins), n), atol=1e-13)
    assert abs(res["resultTypes"][1]["value"] - O.sv_expect_pauli(O.evolve_sv(n, ins), n, "IIZ" + "I" * (n - 3))) < 1e-12
    os.unlink(paths[0])
