"""Reduced density matrices: bsim_partial_trace on the device and the host
mirror of partial_trace / calculate(DensityMatrix) above it.

The host logic (special cases, unsorted-target reindexing, full-qubit
permutation) is checked on the CPU against the oracle's loop-for-loop
restatement with the device reduction replaced by the oracle's sorted
reduction; the device reduction is checked on the GPU against the oracle
(1e-12) and, at 24 qubits, against a numpy Gram product.
"""
import json
import sys
from types import SimpleNamespace

import numpy as np
import pytest

import oracle as O
from bsim import _lib as L
from bsim.results import DensityMatrix, calculate
from bsim.simulators import DensityMatrixSimulator, StateVectorSimulator, _Simulator

from conftest import GOLDEN

TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rand_state(n, rng):
    psi = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    return psi / np.linalg.norm(psi)


def _rand_rho(n, rng):
    a = rng.normal(size=(2 ** n, 2 ** n)) + 1j * rng.normal(size=(2 ** n, 2 ** n))
    r = a @ a.conj().T
    return r / np.trace(r)


TARGET_SETS = [[], [0], [2], [1, 3], [3, 1], [0, 2, 1], [2, 0, 4], [4, 3, 2, 1, 0], [0, 1, 2, 3, 4], [1, 4]]


def _fake(n, kind, state):
    """A stand-in simulator whose device reduction is the oracle's."""
    dm = kind == L.BSIM_DM

    def reduced(t):
        if not t:
            v = np.trace(state) if dm else np.sum(np.abs(state) ** 2)
            return np.array([[v]])
        return O.partial_trace(state, n, t, dm=dm)

    return SimpleNamespace(qubit_count=n, kind=kind, _reduced=reduced,
                           density_matrix=lambda: state if dm else np.outer(state, np.conj(state)))


@pytest.mark.parametrize("dm", [False, True])
def test_host_partial_trace_logic(dm):
    n = 5
    rng = np.random.default_rng(4)
    state = _rand_rho(n, rng) if dm else _rand_state(n, rng)
    f = _fake(n, L.BSIM_DM if dm else L.BSIM_SV, state)
    for t in TARGET_SETS:
        got = _Simulator.partial_trace(f, t)
        exp = O.partial_trace(state, n, t, dm=dm)
        np.testing.assert_allclose(got, exp, atol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [3, 5, 8, 11])
def test_partial_trace_sv_vs_oracle(n):
    rng = np.random.default_rng(n)
    psi = _rand_state(n, rng)
    sim = StateVectorSimulator(n, 0)
    sim.upload(psi)
    _say(f"partial trace sv n={n}")
    for t in TARGET_SETS + [list(range(min(n, 8)))]:
        t = [q for q in t if q < n]
        if len(set(t)) != len(t):
            continue
        got = sim.partial_trace(t)
        exp = O.partial_trace(psi, n, t, dm=False)
        np.testing.assert_allclose(got, exp, atol=TOL)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [3, 5, 7])
def test_partial_trace_dm_vs_oracle(n):
    rng = np.random.default_rng(10 + n)
    rho = _rand_rho(n, rng)
    sim = DensityMatrixSimulator(n, 0)
    sim.upload(rho)
    _say(f"partial trace dm n={n}")
    stored = sim.density_matrix()
    np.testing.assert_allclose(stored, rho, atol=1e-15)
    for t in TARGET_SETS:
        t = [q for q in t if q < n]
        if len(set(t)) != len(t):
            continue
        got = sim.partial_trace(t)
        exp = O.partial_trace(rho, n, t, dm=True)
        np.testing.assert_allclose(got, exp, atol=TOL)


@pytest.mark.gpu
def test_density_matrix_result_golden():
    """calculate(DensityMatrix(targets)) on the reference's test state
    (test/test_result_types.jl:44-60, :212-219)."""
    d = json.loads((GOLDEN / "density_matrix_results.json").read_text())
    cm = lambda rows: np.array([[complex(*x) for x in r] for r in rows])
    k = np.arange(16)
    psi = np.sqrt(k / 120.0) * np.where(k < 8, 1.0, 1j)
    sim = StateVectorSimulator(4, 0)
    sim.upload(psi)
    _say("density matrix result golden")
    np.testing.assert_allclose(calculate(DensityMatrix([3]), sim), cm(d["density_matrix_3"]), rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(calculate(DensityMatrix([0, 2, 1]), sim), cm(d["density_matrix_0_2_1"]), rtol=1e-6,
                               atol=1e-8)
    full = np.outer(psi, np.conj(psi))
    np.testing.assert_allclose(calculate(DensityMatrix([0, 1, 2, 3]), sim), full, atol=1e-15)
    np.testing.assert_allclose(calculate(DensityMatrix([3, 2, 1, 0]), sim), full, atol=1e-15)


@pytest.mark.gpu
def test_partial_trace_large_sv():
    n = 24
    rng = np.random.default_rng(24)
    psi = _rand_state(n, rng)
    sim = StateVectorSimulator(n, 0)
    sim.upload(psi)
    _say("partial trace sv n=24")
    for t in ([5, 17], [0, 1, 2, 3, 4, 5, 6, 7], [23, 0, 12]):
        got = sim.partial_trace(sorted(t))
        # numpy Gram product: move the targets to the front (first target = most significant)
        T = psi.reshape([2] * n)
        rest = [q for q in range(n) if q not in sorted(t)]
        M = np.transpose(T, sorted(t) + rest).reshape(2 ** len(t), -1)
        np.testing.assert_allclose(got, M @ M.conj().T, atol=1e-12)
        assert abs(np.trace(got) - 1.0) < 1e-12
