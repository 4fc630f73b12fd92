"""Result wire format on the CPU: IR lowering of result types and
observables (the shapes the reference's own Python-extension tests read
back, test/test_python_ext.jl:70-116), value lowering, the measurement-row
text, file-backed large values and the JSON round trip."""
import json
import os

import numpy as np
import pytest

from bsim import observables as Obs
from bsim import wire
from bsim.results import Amplitude, DensityMatrix, Expectation, Probability, Sample, StateVector, Variance


def test_result_type_ir_forms():
    # test/test_python_ext.jl:77,92-95,112-113
    assert wire.lower_result_type(StateVector()) == {"type": "statevector"}
    assert wire.lower_result_type(Amplitude(["00", "01", "10", "11"])) == {"states": ["00", "01", "10", "11"],
                                                                          "type": "amplitude"}
    assert wire.lower_result_type(Expectation(Obs.Z(), [0])) == {"observable": ["z"], "targets": [0],
                                                                 "type": "expectation"}
    assert wire.lower_result_type(DensityMatrix()) == {"targets": None, "type": "densitymatrix"}
    assert wire.lower_result_type(Probability([3, 1])) == {"targets": [3, 1], "type": "probability"}
    assert wire.lower_result_type(Variance(Obs.TensorProduct([Obs.X(), Obs.Y(), Obs.H()]), [0, 2, 1])) == {
        "observable": ["x", "y", "h"], "targets": [0, 2, 1], "type": "variance"}
    assert wire.lower_result_type(Sample(Obs.I())) == {"observable": ["i"], "targets": None, "type": "sample"}


def test_hermitian_observable_ir():
    m = np.array([[1, 1 - 2j], [1 + 2j, -1]])
    low = wire.lower_observable(Obs.TensorProduct([Obs.Z(), Obs.HermitianObservable(m)]))
    assert low[0] == "z"
    assert low[1] == [[[1.0, 0.0], [1.0, -2.0]], [[1.0, 2.0], [-1.0, 0.0]]]  # rows of [re, im]


def test_value_lowering():
    assert wire.lower_value(0.25) == 0.25
    assert wire.lower_value(1 - 2j) == [1.0, -2.0]
    r = 2 ** -0.5
    assert wire.lower_value({"00": r + 0j, "11": 1j * r}) == {"00": [r, 0.0], "11": [0.0, r]}
    v = wire.lower_value(np.array([1 + 2j, 3 - 4j]))
    np.testing.assert_array_equal(v, [[1, 2], [3, -4]])
    m = wire.lower_value(np.array([[1, 2j], [-2j, 1]]))
    assert m == [[[1.0, 0.0], [0.0, 2.0]], [[0.0, -2.0], [1.0, 0.0]]]
    np.testing.assert_array_equal(wire.lower_value([0.5, 0.25]), [0.5, 0.25])


@pytest.mark.parametrize("shots,m", [(1, 1), (7, 3), (1000, 30), (5, 0), (0, 4)])
def test_measurements_text(shots, m):
    rng = np.random.default_rng(shots + m)
    bits = rng.integers(0, 2, size=(shots, m)).astype(np.int8)
    text = wire.measurements_json(bits)
    assert text == json.dumps(bits.tolist(), separators=(",", ":"))
    assert json.loads(text) == bits.tolist()


def _synthetic_result(big_n):
    probs = np.linspace(0, 1, big_n)
    sv = (np.arange(big_n) + 1j * np.arange(big_n)[::-1]).astype(np.complex128)
    return {"resultTypes": [
        {"type": {"type": "expectation"}, "value": 0.5},
        {"type": {"type": "probability"}, "value": wire.lower_value(probs)},
        {"type": {"type": "statevector"}, "value": wire.lower_value(sv)},
        {"type": {"type": "probability"}, "value": wire.lower_value(probs[:8])},
    ]}, probs, sv


def test_large_values_go_to_files(tmp_path):
    big = 2 ** 18  # 2 MiB of float64, 4 MiB of complex
    res, probs, sv = _synthetic_result(big)
    paths, lens = wire.mmap_large_result_values(res, str(tmp_path))
    assert lens == (big, big)
    assert res["resultTypes"][1]["value"] == [] and res["resultTypes"][2]["value"] == []
    assert res["resultTypes"][0]["value"] == 0.5 and len(res["resultTypes"][3]["value"]) == 8
    np.testing.assert_array_equal(np.fromfile(paths[0], dtype="<f8"), probs)
    raw = np.fromfile(paths[1], dtype="<f8").reshape(-1, 2)
    np.testing.assert_array_equal(raw[:, 0] + 1j * raw[:, 1], sv)
    back = wire.load_result(wire.to_json(res), paths, lens)
    np.testing.assert_array_equal(back["resultTypes"][1]["value"], probs)
    np.testing.assert_array_equal(back["resultTypes"][2]["value"][:, 1], sv.imag)
    for p in paths:
        os.unlink(p)


def test_small_values_stay_inline():
    res, probs, _ = _synthetic_result(1024)  # 8 KiB: below 2^20 bytes
    assert wire.mmap_large_result_values(res) == (None, None)
    back = json.loads(wire.to_json(res))
    np.testing.assert_array_equal(back["resultTypes"][1]["value"], probs)


def test_json_splices_measurement_text():
    bits = np.array([[0, 1, 1], [1, 0, 0]], dtype=np.int8)
    res = {"measurements": None, "measuredQubits": [0, 1, 2], "resultTypes": []}
    back = json.loads(wire.to_json(res, wire.measurements_json(bits)))
    assert back["measurements"] == [[0, 1, 1], [1, 0, 0]]
    assert back["measuredQubits"] == [0, 1, 2]


def test_create_sim_has_no_state_until_reinit():
    """create_sim (src/BraketSimulator.jl:402-408) makes 0-qubit simulators;
    they hold no device state, and using one before reinit is an error from
    the library, not a silent no-op."""
    from bsim import BsimError
    from bsim import _lib as L

    L.load()
    for sid in ("braket_sv_v2", "braket_dm_v2"):
        sim = wire.create_sim(sid, 100)
        assert sim.qubit_count == 0 and sim.shots == 100 and not sim.handle.value
        with pytest.raises(BsimError):
            L.call("bsim_state_reset", sim.handle)
    with pytest.raises(ValueError):
        wire.create_sim("braket_xx_v2")
