"""Write tests/golden/closed_forms.json: the closed-form expected values of
BraketSimulator.jl's own tests, evaluated with numpy at the angles those
tests use.  The reference is not run: every formula below is the one its
test file states (cited per case), so the fixture pins the custom-gate
matrices and the result-type paths independently of this package.

    python tests/golden/make_closed_forms.py

Op format (as in reference_fixtures.json): [name, args, targets].  Extra
names understood by tests/circuits.py:
  ["GPhaseN", [n, angle], targets]            GPhase{n}(angle)
  ["ControlOf", [inner, bitvals], targets]    Control(inner, bitvals); inner
                                              is [name, args] (recursively)
  ["UnitaryOf", inner, targets]               Unitary(Matrix(matrix_rep(inner)))
Observables: a list of factors, one factor = that observable, more = a
TensorProduct; a factor is "X", "Y", "Z", "H", "I" or ["Herm", matrix]
with complex entries as [re, im].
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "closed_forms.json"
SQ2 = math.sqrt(2.0)


def cx(v):
    return [[float(np.real(z)), float(np.imag(z))] for z in np.asarray(v, dtype=np.complex128).ravel()]


def cmat(m):
    m = np.asarray(m, dtype=np.complex128)
    return [[[float(z.real), float(z.imag)] for z in row] for row in m]


def pauli_eigenvalues(N: int, coeff: float = 1.0):
    """PauliEigenvalues{N} getindex (src/operators.jl:26-50), 1-based i."""

    def get(N, i):
        if N == 1:
            return (coeff, -coeff)[i - 1]
        i_block = (i - 1) // 2
        split = (2 ** (N - 1) - 1) // 2
        if N < 5:
            is_front = (i - 1) % 2 == 0
            ev = coeff if is_front else -coeff
            mi, di = i_block % 2, i_block // 2
            if i_block <= split:
                return -ev if (mi % 2 == 1) ^ (di % 2 == 1) else ev
            mi, di = (i_block - split - 1) % 2, (i_block - split - 1) // 2
            return ev if (mi % 2 == 1) ^ (di % 2 == 1) else -ev
        new_i = i - 2 ** (N - 1) if i > 2 ** (N - 1) else i
        one_down = get(N - 1, new_i)
        return one_down if i_block <= split else -one_down

    return np.array([get(N, i) for i in range(1, 2 ** N + 1)])


def H(q):
    return ["H", [], [q]]


def CNOT(c, t):
    return ["CNot", [], [c, t]]


def RY(a, q):
    return ["Ry", [a], [q]]


def custom_gate_cases():
    """test/test_custom_gates.jl: each case gives its instruction sets
    (raw gate, decomposition, Unitary(matrix_rep(gate))), the state vector
    (checked on the state-vector simulator) and the probabilities (checked
    on both simulators)."""
    cases = []
    phi = 3.56
    c, s = math.cos(phi / 2), math.sin(phi / 2)

    # Double excitation (:4-32)
    de = [H(0), CNOT(0, 1), CNOT(2, 3), CNOT(0, 2), H(3), H(0), CNOT(2, 3), CNOT(0, 1), RY(phi / 8, 1),
          RY(-phi / 8, 0), CNOT(0, 3), H(3), CNOT(3, 1), RY(phi / 8, 1), RY(-phi / 8, 0), CNOT(2, 1), CNOT(2, 0),
          RY(-phi / 8, 1), RY(phi / 8, 0), CNOT(3, 1), H(3), CNOT(0, 3), RY(-phi / 8, 1), RY(phi / 8, 0),
          CNOT(0, 1), CNOT(2, 0), H(0), H(3), CNOT(0, 2), CNOT(2, 3)]
    sv = np.zeros(16, dtype=np.complex128)
    sv[0], sv[3], sv[12] = 1 / SQ2, -s / SQ2, c / SQ2
    cases.append({"cite": "test/test_custom_gates.jl:4-32", "name": "DoubleExcitation", "n": 4,
                  "sets": {"raw": [H(0), CNOT(0, 1), ["DoubleExcitation", [phi], [0, 1, 2, 3]]],
                           "decomp": de,
                           "unitary": [H(0), CNOT(0, 1), ["UnitaryOf", ["DoubleExcitation", [phi]], [0, 1, 2, 3]]]},
                  "state": cx(sv), "probs": list(np.abs(sv) ** 2)})

    # Single excitation (:33-60)
    de = [H(0), H(1), ["Ti", [], [0]], H(0), ["S", [], [0]], ["Ti", [], [1]], ["Si", [], [1]], H(1), CNOT(1, 0),
          ["Rz", [-phi / 2], [0]], RY(phi / 2, 1), CNOT(1, 0), ["Si", [], [0]], H(0), ["T", [], [0]], H(1),
          ["S", [], [1]], ["T", [], [1]]]
    sv = 0.5 * np.array([1, c - s, c + s, 1], dtype=np.complex128)
    cases.append({"cite": "test/test_custom_gates.jl:33-60", "name": "SingleExcitation", "n": 2,
                  "sets": {"raw": [H(0), H(1), ["SingleExcitation", [phi], [0, 1]]], "decomp": de,
                           "unitary": [H(0), H(1), ["UnitaryOf", ["SingleExcitation", [phi]], [0, 1]]]},
                  "state": cx(sv), "probs": list(0.25 * np.array([1, (c - s) ** 2, (c + s) ** 2, 1]))})

    # 3-angle U (:61-88): the test states the literals
    th, ph, lam = 1.34, 2.12, 0.43
    sv = np.array([0.10968333631828209 - 0.12943546335661305j, 0.10968333631828209 - 0.12943546335661305j,
                   -0.4873868533239797 + 0.483394333610458j, -0.4873868533239797 + 0.483394333610458j])
    cases.append({"cite": "test/test_custom_gates.jl:61-88", "name": "U", "n": 2,
                  "sets": {"raw": [H(0), H(1), ["U", [th, ph, lam], [0]]],
                           "unitary": [H(0), H(1), ["UnitaryOf", ["U", [th, ph, lam]], [0]]]},
                  "state": cx(sv),
                  "probs": [0.0287839734402505, 0.0287839734402505, 0.47121602655974926, 0.47121602655974926]})

    # GPhase{nq} (:89-110)
    ph = 2.12
    for nq in range(1, 5):
        sv = np.exp(1j * ph) / math.sqrt(2 ** nq) * np.ones(2 ** nq)
        hs = [H(q) for q in range(nq)]
        cases.append({"cite": "test/test_custom_gates.jl:89-110", "name": f"GPhase{nq}", "n": nq,
                      "sets": {"raw": hs + [["GPhaseN", [nq, ph], list(range(nq))]],
                               "unitary": hs + [["UnitaryOf", ["GPhaseN", [nq, ph]], list(range(nq))]]},
                      "state": cx(sv), "probs": [1.0 / 2 ** nq] * 2 ** nq})

    # MultiRZ (:111-127)
    for nq in range(1, 7):
        sv = math.sqrt(1 / 2 ** nq) * np.exp(-1j * ph / 2 * pauli_eigenvalues(nq))
        cases.append({"cite": "test/test_custom_gates.jl:111-127", "name": f"MultiRZ{nq}", "n": nq,
                      "sets": {"raw": [H(q) for q in range(nq)] + [["MultiRZ", [ph], list(range(nq))]]},
                      "state": cx(sv), "probs": [1.0 / 2 ** nq] * 2 ** nq})

    # Control (:128-178): Control(X, (1,)) and Control(Control(X, (1,)), (1,)) on GHZ preparations
    cx_spec = ["Control", [["X", []], [1]]]
    ccx_spec = ["Control", [cx_spec, [1]]]
    for nq, g, builtin in ((2, cx_spec, "CNot"), (3, ccx_spec, "CCNot")):
        v = np.zeros(2 ** nq)
        v[0] = v[-1] = 1
        prep = [H(0)] + [CNOT(q, q + 1) for q in range(nq - 2)]
        cases.append({"cite": "test/test_custom_gates.jl:128-158", "name": f"Control{nq}", "n": nq,
                      "sets": {"control": prep + [["ControlOf", g[1], list(range(nq))]],
                               "builtin": prep + [[builtin, [], list(range(nq))]]},
                      "state": cx(v / SQ2), "probs": list(v / 2)})
    for nq in (1, 2):
        inner = ["GPhaseN", [nq, ph]]
        sv = 1 / math.sqrt(2 ** nq) * np.ones(2 ** nq, dtype=np.complex128)
        sv[0] *= np.exp(1j * ph)
        hs = [H(q) for q in range(nq)]
        cases.append({"cite": "test/test_custom_gates.jl:159-178", "name": f"ControlGPhase{nq}", "n": nq,
                      "sets": {"raw": hs + [["ControlOf", [inner, [0] * nq], list(range(nq))]],
                               "unitary": hs + [["UnitaryOf", ["Control", [inner, [0] * nq]], list(range(nq))]]},
                      "state": cx(sv), "probs": [1.0 / 2 ** nq] * 2 ** nq})

    # Single excitation plus / minus (:180-227)
    for name, sign, cite in (("SingleExcitationPlus", 1, "180-203"), ("SingleExcitationMinus", -1, "204-227")):
        de = [H(0), H(1), ["X", [], [0]], ["X", [], [1]], ["CPhaseShift", [sign * phi / 2], [1, 0]], ["X", [], [0]],
              ["X", [], [1]], ["CPhaseShift", [sign * phi / 2], [0, 1]], CNOT(0, 1), RY(phi / 2, 0), CNOT(1, 0),
              RY(-phi / 2, 0), CNOT(1, 0), CNOT(0, 1)]
        e = np.exp(sign * 1j * phi / 2)
        sv = 0.5 * np.array([e, c - s, c + s, e])
        cases.append({"cite": f"test/test_custom_gates.jl:{cite}", "name": name, "n": 2,
                      "sets": {"raw": [H(0), H(1), [name, [phi], [0, 1]]], "decomp": de,
                               "unitary": [H(0), H(1), ["UnitaryOf", [name, [phi]], [0, 1]]]},
                      "state": cx(sv), "probs": list(0.25 * np.array([1, (c - s) ** 2, (c + s) ** 2, 1]))})

    # Double excitation minus / plus (:228-269)
    for name, sign, cite in (("DoubleExcitationMinus", -1, "228-248"), ("DoubleExcitationPlus", 1, "249-269")):
        e = np.exp(sign * 1j * phi / 2.0)
        sv = 0.5 * np.array([e, 0, 0, -s, e, 0, 0, 0, e, 0, 0, 0, c, 0, 0, 0])
        cases.append({"cite": f"test/test_custom_gates.jl:{cite}", "name": name, "n": 4,
                      "sets": {"raw": [H(0), H(1), [name, [phi], [0, 1, 2, 3]]],
                               "unitary": [H(0), H(1), ["UnitaryOf", [name, [phi]], [0, 1, 2, 3]]]},
                      "state": cx(sv), "probs": list(np.abs(sv) ** 2)})

    # FermionicSWAP (:270-293); the Unitary form is applied on [1, 0]
    e = np.exp(1j * phi / 2.0)
    sv = 0.5 * np.array([1, e * c - 1j * e * s, -1j * e * s + e * c, np.exp(1j * phi)])
    cases.append({"cite": "test/test_custom_gates.jl:270-293", "name": "FermionicSWAP", "n": 2,
                  "sets": {"raw": [H(0), H(1), ["FermionicSWAP", [phi], [0, 1]]],
                           "unitary": [H(0), H(1), ["UnitaryOf", ["FermionicSWAP", [phi]], [1, 0]]]},
                  "state": cx(sv), "probs": list(np.abs(sv) ** 2)})
    return cases


def matrix_cases():
    """matrix_rep of Control gates (test/test_custom_gates.jl:128-139)."""
    xm = [[0, 1], [1, 0]]
    cnot = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]
    cx_spec = ["Control", [["X", []], [1]]]
    return [
        {"cite": "test/test_custom_gates.jl:129-130", "gate": ["Control", [["X", []], []]], "matrix": cmat(xm)},
        {"cite": "test/test_custom_gates.jl:132-133", "gate": cx_spec, "matrix": cmat(cnot)},
        {"cite": "test/test_custom_gates.jl:135-136", "gate": ["Control", [cx_spec, []]], "matrix": cmat(cnot)},
    ]


def integration_cases():
    """test/test_braket_integration.jl, the shots = 0 result-type checks of
    the 3-qubit Rx/CNOT circuit and the Bell circuits."""
    th, ph, vp = 0.432, 0.123, -0.543
    three = [["Rx", [th], [0]], ["Rx", [ph], [1]], ["Rx", [vp], [2]], CNOT(0, 1), CNOT(1, 2)]
    bell = [H(0), CNOT(0, 1)]
    X = np.array([[0, 1], [1, 0]])
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1.0, -1.0])
    cases = []

    def herm(m):
        return ["Herm", cmat(m)]

    def add(cite, n, ops, obs, targets, mean=None, var=None):
        cases.append({"cite": cite, "n": n, "ops": ops, "observable": obs, "targets": targets,
                      "mean": mean, "var": var})

    mean_xy = math.sin(th) * math.sin(ph) * math.sin(vp)
    var_xy = (8 * math.sin(th) ** 2 * math.cos(2 * vp) * math.sin(ph) ** 2 - math.cos(2 * (th - ph))
              - math.cos(2 * (th + ph)) + 2 * math.cos(2 * th) + 2 * math.cos(2 * ph) + 14) / 16
    for obs in (["X", "Y"], [herm(np.kron(X, Y))]):
        add("test/test_braket_integration.jl:204-238", 3, three, obs, [0, 2], mean_xy, var_xy)

    mean_zhy = -(math.cos(vp) * math.sin(ph) + math.sin(vp) * math.cos(th)) / SQ2
    var_zhy = (3 + math.cos(2 * ph) * math.cos(vp) ** 2 - math.cos(2 * th) * math.sin(vp) ** 2
               - 2 * math.cos(th) * math.sin(ph) * math.sin(2 * vp)) / 4
    add("test/test_braket_integration.jl:239-272", 3, three, ["Z", "H", "Y"], [0, 1, 2], mean_zhy, var_zhy)

    for obs in (["Z", "Z"], [herm(np.kron(Z, Z))]):
        add("test/test_braket_integration.jl:273-301", 3, three, obs, [0, 2], 0.849694136476246, 0.27801987443788634)

    ho = np.array([[-6, 2 + 1j, -3, -5 + 2j], [2 - 1j, 0, 2 - 1j, -5 + 4j], [-3, 2 + 1j, 0, -4 + 3j],
                   [-5 - 2j, -5 - 4j, -4 - 3j, -6]])
    ho2 = np.array([[1, 2], [2, 4]])
    ho3 = np.array([[-6, 2 + 1j], [2 - 1j, 0]])
    ho4 = np.kron(np.eye(2), ho3)
    meani, meany, meanh, meanii = -5.7267957792059345, 1.4499810303182408, -4.30215023196904, -5.78059066879935
    meanz = 0.5 * (-6 * math.cos(th) * (math.cos(vp) + 1) - 2 * math.sin(vp) * (math.cos(th) + math.sin(ph)
                   - 2 * math.cos(ph)) + 3 * math.cos(vp) * math.sin(ph) + math.sin(ph))
    vari, vary, varh, varii = 43.33800156673375, 74.03174647518193, 370.71292282796804, 6.268315532585994
    c2p, s2p = math.cos(2 * ph), math.sin(2 * ph)
    varz = (1057 - c2p + 12 * (27 + c2p) * math.cos(vp)
            - 2 * math.cos(2 * vp) * math.sin(ph) * (16 * math.cos(ph) + 21 * math.sin(ph)) + 16 * s2p
            - 8 * (-17 + c2p + 2 * s2p) * math.sin(vp)
            - 8 * math.cos(2 * th) * (3 + 3 * math.cos(vp) + math.sin(vp)) ** 2
            - 24 * math.cos(ph) * (math.cos(ph) + 2 * math.sin(ph)) * math.sin(2 * vp)
            - 8 * math.cos(th) * (4 * math.cos(ph) * (4 + 8 * math.cos(vp) + math.cos(2 * vp)
                                                       - (1 + 6 * math.cos(vp)) * math.sin(vp))
                                  + math.sin(ph) * (15 + 8 * math.cos(vp) - 11 * math.cos(2 * vp)
                                                    + 42 * math.sin(vp) + 3 * math.sin(2 * vp)))) / 16
    cite = "test/test_braket_integration.jl:302-394"
    for obs, m, v in ((["I", herm(ho)], meani, vari), (["Y", herm(ho)], meany, vary),
                      (["Z", herm(ho)], meanz, varz), ([herm(ho2), herm(ho)], meanh, varh),
                      ([herm(np.kron(ho2, ho))], meanh, varh), (["I", "I", herm(ho3)], meanii, varii),
                      (["I", herm(ho4)], meanii, varii)):
        add(cite, 3, three, obs, [0, 1, 2], m, v)
    cite = "test/test_braket_integration.jl:395-456"
    for obs, t, m, v in (([herm(ho)], [1, 2], meani, vari), ([herm(np.kron(ho2, ho))], [0, 1, 2], meanh, varh),
                         ([herm(ho3)], [2], meanii, varii), ([herm(ho4)], [1, 2], meanii, varii)):
        add(cite, 3, three, obs, t, m, v)

    # all selected (:457-510): Variance(ho) over every qubit, Expectation(ho, 0)
    a = 0.543
    hm = np.array([[1, 2j], [-2j, 0]])
    var_ = 0.25 * (math.sin(a) - 4 * math.cos(a)) ** 2
    ops2 = [["Rx", [a], [0]], ["Rx", [a], [1]]]
    add("test/test_braket_integration.jl:457-510", 2, ops2, [herm(hm)], [], None, [var_, var_])
    add("test/test_braket_integration.jl:457-510", 2, ops2, [herm(hm)], [0],
        2 * math.sin(a) + 0.5 * math.cos(a) + 0.5, None)

    # noncommuting (:511-553)
    cite = "test/test_braket_integration.jl:511-553"
    add(cite, 3, three, ["X", "Y"], [0, 2], mean_xy, var_xy)
    add(cite, 3, three, ["Z", "Z"], [0, 2], 0.849694136476246, None)
    add(cite, 3, three, ["Y", herm(ho)], [0, 1, 2], 1.4499810303182408, None)
    # flipped targets (:554-566)
    cite = "test/test_braket_integration.jl:554-566"
    add(cite, 2, bell, ["H", "X"], [0, 1], SQ2 / 2, None)
    add(cite, 2, bell, ["H", "X"], [1, 0], SQ2 / 2, None)
    # all noncommuting (:567-579)
    cite = "test/test_braket_integration.jl:554-566"
    add(cite, 2, bell, [herm(hm)], [], [0.5, 0.5], None)
    add(cite, 2, bell, ["X"], [], [0.0, 0.0], None)
    # observable not in instructions (:580-595): a 4-qubit register
    cite = "test/test_braket_integration.jl:580-595"
    add(cite, 4, bell, ["X"], [2], 0.0, None)
    add(cite, 4, bell, ["Y"], [3], None, 1.0)
    return cases


def main():
    pe = [pauli_eigenvalues(n) for n in range(1, 7)]
    for n, p in enumerate(pe, start=1):  # the reference's table is Z^(x)n's spectrum
        assert np.array_equal(p, [(-1.0) ** bin(i).count("1") for i in range(2 ** n)]), n
    doc = {
        "about": "Closed forms of BraketSimulator.jl's own tests (custom gates: test/test_custom_gates.jl; "
                 "result types: test/test_braket_integration.jl, shots = 0), evaluated with numpy by "
                 "tests/golden/make_closed_forms.py at the angles of those tests.  Complex numbers are "
                 "[re, im].  Integration entries: mean / var are numbers, lists (one per target when the "
                 "targets are empty) or null (not checked by that reference test).",
        "custom_gates": custom_gate_cases(),
        "matrices": matrix_cases(),
        "integration": integration_cases(),
    }
    OUT.write_text(json.dumps(doc, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
