"""GPU parity against the reference's own closed forms and the oracle:
custom gates (test/test_custom_gates.jl) on both simulators in every
evolve mode, the result-type closed forms of
test/test_braket_integration.jl (Expectation and Variance, shots = 0), the
DoubleExcitation kernel (src/custom_gates.jl:390-416) on low, high and
scattered targets, and Variance (src/result_types.jl:168-194,223-238) of
Pauli words, Hermitian matrices and tensor products on random states.

Bar: closed forms to 1e-12 (amplitudes) / 1e-11 relative (the integration
literals carry 15-16 digits); the oracle to 1e-10 absolute.
"""
import json
import math
import sys

import numpy as np
import pytest

import oracle as O
from bsim import gates as G
from bsim import observables as Ob
from bsim.results import Expectation, Variance, calculate
from bsim.simulators import DensityMatrixSimulator, Instruction, StateVectorSimulator

from circuits import fixture_observable, fixture_ops, random_circuit, random_noise, random_unitary
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

CF = json.loads((GOLDEN / "closed_forms.json").read_text())
TOL = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cx(rows):
    return np.array([complex(a, b) for a, b in rows])


def _cases():
    for case in CF["custom_gates"]:
        for label, ops in case["sets"].items():
            yield pytest.param(case, ops, id=f"{case['name']}-{label}")


@pytest.mark.parametrize("mode", ["fused", "per_op", "apply_gate"])
@pytest.mark.parametrize("case,ops", list(_cases()))
def test_custom_gate_closed_form_sv(case, ops, mode):
    n = case["n"]
    sim = StateVectorSimulator(n, 0)
    ins = fixture_ops(case, ops)
    if mode == "apply_gate":
        for i in ins:
            sim.apply_gate(i.operator, *i.target)
    else:
        sim.evolve(ins, mode=0 if mode == "fused" else 1)
    np.testing.assert_allclose(sim.state_vector(), _cx(case["state"]), atol=1e-12, err_msg=case["cite"])
    np.testing.assert_allclose(sim.probabilities(), case["probs"], atol=1e-12, err_msg=case["cite"])


@pytest.mark.parametrize("noise_mode", ["reference", "physical"])
@pytest.mark.parametrize("mode", ["fused", "per_op", "apply_gate"])
@pytest.mark.parametrize("case,ops", list(_cases()))
def test_custom_gate_closed_form_dm(case, ops, mode, noise_mode):
    n = case["n"]
    sim = DensityMatrixSimulator(n, 0, noise_mode=noise_mode)
    ins = fixture_ops(case, ops)
    if mode == "apply_gate":
        for i in ins:
            sim.apply_gate(i.operator, *i.target)
    else:
        sim.evolve(ins, mode=0 if mode == "fused" else 1)
    np.testing.assert_allclose(sim.probabilities(), case["probs"], atol=1e-12, err_msg=case["cite"])
    psi = _cx(case["state"])
    np.testing.assert_allclose(sim.density_matrix(), np.outer(np.conj(psi), psi), atol=1e-12, err_msg=case["cite"])


def _factors(spec):
    return [f if isinstance(f, str) else np.array([[complex(*z) for z in row] for row in f[1]]) for f in spec]


def _close(got, exp, what, rtol=1e-11):
    got = np.atleast_1d(np.asarray(got, dtype=float))
    exp = np.atleast_1d(np.asarray(exp, dtype=float))
    np.testing.assert_allclose(got, exp, rtol=rtol, atol=rtol, err_msg=what)


@pytest.mark.parametrize("dm", [False, True], ids=["sv", "dm"])
@pytest.mark.parametrize("i", range(len(CF["integration"])))
def test_integration_closed_forms(i, dm):
    case = CF["integration"][i]
    n = case["n"]
    ins = fixture_ops(case)
    sim = DensityMatrixSimulator(n, 0) if dm else StateVectorSimulator(n, 0)
    sim.evolve(ins)
    obs = fixture_observable(case["observable"])
    st = O.evolve_dm(n, ins) if dm else O.evolve_sv(n, ins)
    fs = _factors(case["observable"])
    if case["mean"] is not None:
        got = calculate(Expectation(obs, list(case["targets"])), sim)
        _close(got, case["mean"], case["cite"])
        _close(got, O.calculate_expectation(st, n, fs, case["targets"], dm), "oracle", TOL)
    if case["var"] is not None:
        got = calculate(Variance(obs, list(case["targets"])), sim)
        _close(got, case["var"], case["cite"])
        _close(got, O.calculate_variance(st, n, fs, case["targets"], dm), "oracle", TOL)


def _random_states(n, dm, seed):
    rng = np.random.default_rng(seed)
    if dm:
        ops = []
        for _ in range(6):
            ops += random_circuit(n, 3, rng)
            ops.append(random_noise(n, rng))
        sim = DensityMatrixSimulator(n, 0)
        sim.evolve(ops)
        return sim, O.evolve_dm(n, ops), rng
    ops = random_circuit(n, 60, rng)
    sim = StateVectorSimulator(n, 0)
    sim.evolve(ops)
    return sim, O.evolve_sv(n, ops), rng


def _herm(rng, k):
    a = rng.normal(size=(2 ** k, 2 ** k)) + 1j * rng.normal(size=(2 ** k, 2 ** k))
    return a + a.conj().T


@pytest.mark.parametrize("dm", [False, True], ids=["sv", "dm"])
def test_variance_vs_oracle(dm):
    """Variance and Expectation of Pauli words, the H observable, Hermitian
    matrices on 1-3 targets in given (unsorted) order and tensor products
    mixing them, against the oracle's factor-by-factor restatement."""
    n = 6 if dm else 9
    sim, st, rng = _random_states(n, dm, 11 + dm)
    h1, h2, h3 = _herm(rng, 1), _herm(rng, 2), _herm(rng, 3)
    cases = [
        (["Z"], [2]), (["Y"], []), (["H"], [1]), (["X", "Y", "Z"], [0, 4, 2]), (["H", "X"], [3, 1]),
        ([h1], [5]), ([h1], []), ([h2], [1, 4]), ([h2], [4, 1]), ([h3], [0, 2, 5]), ([h3], [5, 0, 2]),
        (["X", h2], [0, 5, 2]), ([h2, "Z"], [3, 1, 0]), ([h1, h2], [2, 5, 4]), (["I", h3], [1, 4, 0, 3]),
    ]
    for fs, t in cases:
        spec = [f if isinstance(f, str) else ["Herm", [[[z.real, z.imag] for z in row] for row in f]] for f in fs]
        obs = fixture_observable(spec)
        _say(f"variance {'dm' if dm else 'sv'} {[f if isinstance(f, str) else f.shape[0] for f in fs]} on {t}")
        _close(calculate(Expectation(obs, list(t)), sim), O.calculate_expectation(st, n, fs, t, dm),
               f"expectation {t}", TOL)
        _close(calculate(Variance(obs, list(t)), sim), O.calculate_variance(st, n, fs, t, dm), f"variance {t}", TOL)


@pytest.mark.parametrize("n", [6, 12, 17])
def test_double_excitation_kernel(n):
    """bsim_apply_double_excitation (apply_gate) and the evolve paths against
    the oracle's kernel on low, high and scattered targets, with pow."""
    rng = np.random.default_rng(n)
    base = random_circuit(n, 30, rng)
    for tg in ([0, 1, 2, 3], [n - 1, n - 2, n - 3, n - 4], [n - 2, 0, n // 2, 1], [2, n - 1, 1, n // 2 + 1]):
        ins = base + [Instruction(G.DoubleExcitation(0.61), tg), Instruction(G.H(), [tg[0]]),
                      Instruction(G.DoubleExcitation(-1.37, pow_exponent=2.5), list(reversed(tg)))]
        exp = O.evolve_sv(n, ins)
        _say(f"double excitation n={n} targets={tg}")
        for mode in ("apply_gate", 0, 1):
            sim = StateVectorSimulator(n, 0)
            if mode == "apply_gate":
                for i in ins:
                    sim.apply_gate(i.operator, *i.target)
            else:
                sim.evolve(ins, mode=mode)
            err = float(np.max(np.abs(sim.state_vector() - exp)))
            assert err < TOL, (mode, tg, err)
        dsim = DensityMatrixSimulator(min(n, 6), 0)
        if n <= 6:
            dsim.evolve(ins)
            assert np.max(np.abs(dsim.density_matrix().ravel(order="F") - O.evolve_dm(n, ins))) < TOL
