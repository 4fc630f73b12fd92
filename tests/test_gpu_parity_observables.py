"""Basis rotations for observables: apply_observables / state_with_observables
on state vectors and density matrices, with the reference's own cases
(test/test_sv_simulator.jl:275-345, test/test_dm_simulator.jl:265-345): the
rotated copy must equal a second simulator evolved with the equivalent
gates, from |0...0> as in the reference and after a random circuit; the
original state is left as it was; a second application and a read before
any application raise."""
import sys

import numpy as np
import pytest

import oracle as O
from bsim import gates as G
from bsim import observables as Obs
from bsim.simulators import DensityMatrixSimulator, Instruction, StateVectorSimulator, diagonalizing_gates

pytestmark = pytest.mark.gpu

CASES = [
    ([(Obs.X(), [0])], [Instruction(G.H(), [0])], 1),
    ([(Obs.Z(), [0])], [], 1),
    ([(Obs.I(), [0])], [], 1),
    ([(Obs.X(), [0]), (Obs.Z(), [3]), (Obs.H(), [2])], [Instruction(G.H(), [0]), Instruction(G.Ry(-np.pi / 4), [2])], 5),
    ([(Obs.TensorProduct([Obs.X(), Obs.Z(), Obs.H(), Obs.I()]), [0, 3, 2, 1])],
     [Instruction(G.H(), [0]), Instruction(G.Ry(-np.pi / 4), [2])], 5),
    ([(Obs.X(), [0, 1])], [Instruction(G.H(), [0]), Instruction(G.H(), [1])], 2),
    ([(Obs.Z(), [0, 1])], [], 2),
    ([(Obs.I(), [0, 1])], [], 2),
    ([(Obs.TensorProduct([Obs.I(), Obs.Z()]), [2, 0])], [], 3),
    ([(Obs.TensorProduct([Obs.X(), Obs.Z()]), [2, 0])], [Instruction(G.H(), [2])], 3),
    ([(Obs.Y(), [1])], [Instruction(G.Unitary(2 ** -0.5 * np.array([[1.0, -1j], [1.0, 1j]])), [1])], 2),
]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _prep(n, seed):
    rng = np.random.default_rng(seed)
    ins = [Instruction(G.U(*rng.uniform(0, 6, 3)), [q]) for q in range(n)]
    ins += [Instruction(G.CNot(), [q, q + 1]) for q in range(n - 1)]
    return ins


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("cls", [StateVectorSimulator, DensityMatrixSimulator])
def test_apply_observables_matches_gates(case, cls):
    obs, gates, n = CASES[case]
    _say(f"observables case {case} {cls.__name__}")
    assert [(type(i.operator).__name__, list(i.target)) for i in diagonalizing_gates_all(obs)] == \
        [(type(i.operator).__name__, list(i.target)) for i in gates]
    for pre in ([], _prep(n, case)):
        a = cls(n, 0)
        a.evolve(pre)
        before = a._result_state().copy()
        a.apply_observables(obs)
        b = cls(n, 0)
        b.evolve(pre + gates)
        np.testing.assert_allclose(a.state_with_observables(), b._result_state(), atol=1e-14)
        np.testing.assert_array_equal(a._result_state(), before)  # the simulator's own state is untouched
        if cls is StateVectorSimulator:
            np.testing.assert_allclose(a.state_with_observables(), O.evolve_sv(n, pre + gates), atol=1e-13)


def diagonalizing_gates_all(obs):
    out = []
    for o, t in obs:
        out += diagonalizing_gates(o, t)
    return out


@pytest.mark.parametrize("cls", [StateVectorSimulator, DensityMatrixSimulator])
def test_apply_observables_errors(cls):
    s = cls(4, 0)
    with pytest.raises(RuntimeError, match="not been applied"):
        s.state_with_observables()
    s.apply_observables([(Obs.X(), [0])])
    with pytest.raises(RuntimeError, match="already been applied"):
        s.apply_observables([(Obs.X(), [0])])


def test_hermitian_observable_rotation():
    """A Hermitian observable on all its targets, and a 2x2 one on several
    targets: the eigenvector matrix as a Unitary gate (src/observables.jl:228-236)."""
    rng = np.random.default_rng(9)
    n = 4
    a = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    h4 = a + a.conj().T
    b = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
    h2 = b + b.conj().T
    pre = _prep(n, 3)
    for obs, t in ((Obs.HermitianObservable(h4), [2, 0]), (Obs.HermitianObservable(h2), [1, 3])):
        _say(f"hermitian rotation {obs.mat.shape} on {t}")
        gates = diagonalizing_gates(obs, t)
        s = StateVectorSimulator(n, 0)
        s.evolve(pre)
        s.apply_observables([(obs, t)])
        np.testing.assert_allclose(s.state_with_observables(), O.evolve_sv(n, pre + gates), atol=1e-13)
        vecs = np.linalg.eigh(obs.mat)[1]
        np.testing.assert_allclose(gates[0].operator.matrix_rep(), vecs, atol=1e-15)
