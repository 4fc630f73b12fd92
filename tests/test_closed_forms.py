"""The oracle against the closed forms of the reference's own tests
(tests/golden/closed_forms.json, written by tests/golden/make_closed_forms.py):
custom gates (test/test_custom_gates.jl) through their raw kernels, their
decompositions into standard gates and their Unitary matrices, on both
simulators; the 3-qubit result-type closed forms of
test/test_braket_integration.jl (expectation and variance, shots = 0, state
vector and density matrix); plus numpy brute force for the oracle functions
those do not reach on their own (the DoubleExcitation kernel on scattered
targets, the density-matrix Hermitian observable on unsorted targets) and a
hand-worked alias table.  CPU only."""
import json
import math

import numpy as np
import pytest

import oracle as O
from circuits import fixture_observable, fixture_ops, gate_from_spec
from conftest import GOLDEN

CF = json.loads((GOLDEN / "closed_forms.json").read_text())
TOL = 1e-12


def _cx(rows):
    return np.array([complex(a, b) for a, b in rows])


def _cases():
    for case in CF["custom_gates"]:
        for label, ops in case["sets"].items():
            yield pytest.param(case, ops, id=f"{case['name']}-{label}")


@pytest.mark.parametrize("case,ops", list(_cases()))
def test_custom_gate_closed_form_sv(case, ops):
    n = case["n"]
    psi = O.evolve_sv(n, fixture_ops(case, ops))
    np.testing.assert_allclose(psi, _cx(case["state"]), atol=TOL, err_msg=case["cite"])
    np.testing.assert_allclose(O.sv_probabilities(psi, n), case["probs"], atol=TOL, err_msg=case["cite"])


@pytest.mark.parametrize("mode", ["reference", "physical"])
@pytest.mark.parametrize("case,ops", list(_cases()))
def test_custom_gate_closed_form_dm(case, ops, mode):
    n = case["n"]
    rho = O.evolve_dm(n, fixture_ops(case, ops), mode=mode)
    np.testing.assert_allclose(O.dm_probabilities(rho, n), case["probs"], atol=TOL, err_msg=case["cite"])
    # pure state: the stored matrix is conj(psi) psi^T (rho^T)
    psi = _cx(case["state"])
    np.testing.assert_allclose(rho.reshape((2 ** n, 2 ** n), order="F"), np.outer(np.conj(psi), psi), atol=TOL)


@pytest.mark.parametrize("case", CF["matrices"], ids=lambda c: c["cite"])
def test_control_matrix_rep(case):
    m = np.array([[complex(*z) for z in row] for row in case["matrix"]])
    np.testing.assert_array_equal(gate_from_spec(case["gate"]).matrix_rep(), m)


def _factors(spec):
    return [f if isinstance(f, str) else np.array([[complex(*z) for z in row] for row in f[1]]) for f in spec]


def _state(case, dm):
    n = case["n"]
    ops = fixture_ops(case)
    return O.evolve_dm(n, ops) if dm else O.evolve_sv(n, ops)


def _close(got, exp, what):
    got = np.atleast_1d(np.asarray(got, dtype=float))
    exp = np.atleast_1d(np.asarray(exp, dtype=float))
    # the literals of the reference tests carry 15-16 digits
    np.testing.assert_allclose(got, exp, rtol=1e-11, atol=1e-11, err_msg=what)


@pytest.mark.parametrize("dm", [False, True], ids=["sv", "dm"])
@pytest.mark.parametrize("i", range(len(CF["integration"])))
def test_integration_closed_forms(i, dm):
    case = CF["integration"][i]
    st = _state(case, dm)
    fs = _factors(case["observable"])
    if case["mean"] is not None:
        _close(O.calculate_expectation(st, case["n"], fs, case["targets"], dm), case["mean"], case["cite"])
    if case["var"] is not None:
        _close(O.calculate_variance(st, case["n"], fs, case["targets"], dm), case["var"], case["cite"])


def test_fixture_observable_matrices():
    """The host observables the GPU tests build from the same specs carry the
    matrices of the specs (kron order of the factors)."""
    for case in CF["integration"]:
        obs = fixture_observable(case["observable"])
        m = np.ones((1, 1), dtype=np.complex128)
        for f in _factors(case["observable"]):
            m = np.kron(m, O._STD[f] if isinstance(f, str) else f)
        np.testing.assert_allclose(obs.matrix(), m, atol=0)


# ---- brute force for the oracle's remaining kernels -------------------------

@pytest.mark.parametrize("targets", [[0, 1, 2, 3], [3, 2, 1, 0], [5, 1, 3, 0], [2, 6, 0, 4]])
def test_double_excitation_bruteforce(targets):
    """ref_apply_double_excitation (src/custom_gates.jl:390-416) against the
    16x16 matrix of matrix_rep_raw (src/custom_gates.jl:41-49: identity but
    rows/columns |0011>, |1100> rotated by phi/2; kron order of the targets
    as given) applied by tensor contraction."""
    n = 7
    rng = np.random.default_rng(sum(targets))
    psi = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    phi = 0.731
    c, s = math.cos(phi / 2), math.sin(phi / 2)
    M = np.eye(16, dtype=np.complex128)
    M[3, 3] = M[12, 12] = c
    M[3, 12], M[12, 3] = -s, s
    exp = O.np_apply(psi, n, M, targets)
    got = psi.copy()
    t = np.array(targets, dtype=np.int32)
    O.lib().ref_apply_double_excitation(O._d(got), n, t.ctypes.data_as(O._ip), c, s)
    np.testing.assert_allclose(got, exp, atol=1e-13)


@pytest.mark.parametrize("targets", [[1], [0, 2], [2, 0], [3, 1, 2], [0, 3, 1]])
def test_dm_hermitian_bruteforce(targets):
    """ref_dm_apply_hermitian (src/dm_simulator.jl:190-224) then the real
    trace equals tr(O rho) for the physical rho, with O's kron order the
    ascending target qubits (the kernel sorts the endian targets), for any
    given target order; the stored matrix is rho^T."""
    n = 4
    N = 2 ** n
    rng = np.random.default_rng(len(targets) * 10 + targets[0])
    A = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    rho = A @ A.conj().T
    rho /= np.trace(rho)
    k = len(targets)
    B = rng.normal(size=(2 ** k, 2 ** k)) + 1j * rng.normal(size=(2 ** k, 2 ** k))
    Hm = B + B.conj().T
    full = np.zeros((N, N), dtype=np.complex128)
    for col in range(N):
        full[:, col] = O.np_apply(np.eye(N, dtype=np.complex128)[:, col].copy(), n, Hm, sorted(targets))
    exp = float(np.real(np.trace(full @ rho)))
    stored = np.ascontiguousarray(rho.T.ravel(order="F"))
    got = O.dm_expect_matrix(stored, n, Hm, targets)
    assert abs(got - exp) < 1e-12
    # and the expectation functions used by the integration checks agree
    assert abs(O.expect_factors(stored, n, [Hm], targets, True) - exp) < 1e-12


def test_alias_table_hand_worked():
    """make_alias_table! (src/result_types.jl:8-54) on p = [0.05, 0.3, 0.05,
    0.4, 0.1, 0.1] (n = 6, wsum = 1), worked by hand with exact binary
    fractions in mind (tolerance 1e-15).

    ap = 6 p = [0.3, 1.8, 0.3, 2.4, 0.6, 0.6].  Classification in index
    order: larges = [2, 4] (1-based), smalls = [1, 3, 5, 6].  The pairing
    pops the LAST small and the LAST large each time:
      s = 6, l = 4: alias[6] = 4, ap[4] = (2.4 - 1) + 0.6 = 2.0  (still large)
      s = 5, l = 4: alias[5] = 4, ap[4] = (2.0 - 1) + 0.6 = 1.6  (still large)
      s = 3, l = 4: alias[3] = 4, ap[4] = (1.6 - 1) + 0.3 = 0.9  -> small
      s = 4, l = 2: alias[4] = 2, ap[2] = (1.8 - 1) + 0.9 = 1.7  (still large)
      s = 1, l = 2: alias[1] = 2, ap[2] = (1.7 - 1) + 0.3 = 1.0  -> pushed to smalls
    and the loop ends (no large left); the final loop sets ap of the
    remaining small (2) to 1.0.  alias[2] is never written (0, i.e. -1 after
    the shift to 0-based); 0-based aliases: [1, -1, 3, 1, 3, 3].
    A FIFO pairing (first small with first large) would give alias[1] = 2
    and alias[3] = 2 instead, so the table pins the LIFO order."""
    p = np.array([0.05, 0.3, 0.05, 0.4, 0.1, 0.1])
    ap, alias = O.alias_reference(p, 1.0)
    np.testing.assert_allclose(ap, [0.3, 1.0, 0.3, 0.9, 0.6, 0.6], atol=1e-15)
    # entries never paired keep alias 0 (1-based), i.e. -1 after the shift
    np.testing.assert_array_equal(alias, [1, -1, 3, 1, 3, 3])
