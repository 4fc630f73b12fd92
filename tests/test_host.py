"""Host-side logic and the C ABI's shape: no GPU work is issued here."""
import ctypes
import math
import re

import numpy as np
import pytest

from conftest import ROOT
import bsim
from bsim import _lib as L
from bsim import gates as G
from bsim import noises as Nz
from bsim.gates import kernel_call
from bsim.simulators import Instruction, lower_instructions
from bsim import workloads as W


def _declared_symbols():
    text = (ROOT / "include" / "bsim.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bsim_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(L.LIB_PATH))
    declared = _declared_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # the ctypes table covers the same set
    assert sorted(L.exported_symbols()) == declared


def test_version_and_error_string_without_gpu():
    lib = L.load()
    assert lib.bsim_version() == 10100
    buf = ctypes.create_string_buffer(64)
    lib.bsim_last_error(buf, 64)


def test_struct_layout_matches_header():
    # bsim_op: 4 ints, 3 x 8 ints, pointer, int + 3 ints, 3 pointers
    assert ctypes.sizeof(L.Op) == 4 * 4 + 3 * 8 * 4 + 8 + 4 * 4 + 3 * 8


def test_gate_matrices_unitary_and_pow_rules():
    # test/test_gate_operations.jl:40-221: inv(g) * g == I and g^p consistency
    rng = np.random.default_rng(0)
    for name, cls in G.GATES.items():
        nang = {"U": 3, "MS": 3, "PRx": 2}.get(name, 0)
        try:
            g = cls(*rng.uniform(-2, 2, nang)) if nang else cls()
        except ValueError:
            g = cls(float(rng.uniform(-2, 2)))
        m = g.matrix_rep()
        np.testing.assert_allclose(m @ m.conj().T, np.eye(m.shape[0]), atol=1e-12, err_msg=name)
        gi = type(g).__new__(type(g))
        gi.__dict__.update(g.__dict__)
        gi.pow_exponent = -1.0
        np.testing.assert_allclose(gi.matrix_rep() @ m, np.eye(m.shape[0]), atol=1e-10, err_msg=name)
        g2 = type(g).__new__(type(g))
        g2.__dict__.update(g.__dict__)
        g2.pow_exponent = 2.0
        np.testing.assert_allclose(g2.matrix_rep(), m @ m, atol=1e-10, err_msg=name)


def test_dispatch_conventions():
    # 2q dense: first target is matrix bit 0
    c = kernel_call(G.Swap(), [3, 1])
    assert c.targets == [3, 1] and not c.controls
    # CNot -> controlled X with control = first target
    c = kernel_call(G.CNot(), [2, 5])
    assert c.controls == [2] and c.targets == [5] and c.cvals == [1]
    np.testing.assert_array_equal(c.matrix, [[0, 1], [1, 0]])
    # CCNot: two controls
    c = kernel_call(G.CCNot(), [0, 1, 2])
    assert c.controls == [0, 1] and c.targets == [2]
    # CSwap: one control, two targets in given order
    c = kernel_call(G.CSwap(), [2, 0, 1])
    assert c.controls == [2] and c.targets == [0, 1]
    # k-qubit fallback: matrix bit j is the j-th smallest endian target
    c = kernel_call(G.Unitary(np.eye(8)), [0, 2, 1])
    assert c.targets == [2, 1, 0]
    # Control bit values
    c = kernel_call(G.Control(G.X(), (0,)), [1, 0])
    assert c.controls == [1] and c.cvals == [0]


def test_lowering_builds_ops():
    ins = [Instruction(G.H(), [0]), Instruction(G.CNot(), [0, 1]), Instruction(G.CPhaseShift(0.3), [1, 2]),
           Instruction(G.Unitary(np.eye(8)), [0, 1, 2])]
    ol = lower_instructions(ins, L.BSIM_SV)
    assert ol.n == 4
    assert ol.arr[0].type == L.OP_GATE and ol.arr[0].n_targets == 1
    assert ol.arr[1].n_controls == 1 and ol.arr[1].controls[0] == 0 and ol.arr[1].targets[0] == 1
    assert ol.arr[2].type == L.OP_GATE  # diagonal matrices are detected by the library
    assert ol.arr[3].n_targets == 3 and list(ol.arr[3].targets)[:3] == [2, 1, 0]


def test_plan_stats_brickwall_fuses():
    """The planner (host C++) packs the 30-qubit brick-wall into far fewer
    HBM passes than gates."""
    ins = W.brickwall(30, 20, seed=0)
    ol = lower_instructions(ins, L.BSIM_SV)
    stats = (ctypes.c_int64 * 4)()
    L.call("bsim_plan_stats", 30, 0, ol.ptr, ol.n, stats)
    passes, singles, sweeps, fused = list(stats)
    assert singles == 0
    assert fused == len(ins)
    assert passes < len(ins) / 10, (passes, len(ins))


def test_plan_stats_dm_and_kq():
    ins = W.brickwall(6, 4, seed=1, noise=True)
    ol = lower_instructions(ins, L.BSIM_DM)
    stats = (ctypes.c_int64 * 4)()
    L.call("bsim_plan_stats", 12, 1, ol.ptr, ol.n, stats)
    assert stats[1] == 0  # 1- and 2-qubit reference Kraus channels run inside the tile passes
    assert stats[3] == ol.n + sum(1 for i in ins if isinstance(i.operator, G.Gate))  # DM gates lower to 2 ops
    # a 3-qubit reference Kraus channel keeps its own kernel
    k3 = Instruction(Nz.Kraus([np.eye(8)]), [0, 2, 4])
    ol3 = lower_instructions(ins[:5] + [k3], L.BSIM_DM)
    L.call("bsim_plan_stats", 12, 1, ol3.ptr, ol3.n, stats)
    assert stats[1] == 1


def test_bad_op_rejected_without_gpu():
    ins = [Instruction(G.H(), [5])]
    ol = lower_instructions(ins, L.BSIM_SV)
    stats = (ctypes.c_int64 * 4)()
    rc = L.load().bsim_plan_stats(3, 0, ol.ptr, ol.n, stats)
    assert rc == -1 and "range" in L.last_error()


def test_derivative_matrices_accuracy():
    th = 0.37
    d = G.derivative_matrices(G.Ry(th), [0])[0]
    exact = 0.5 * np.array([[-math.sin(th / 2), -math.cos(th / 2)], [math.cos(th / 2), -math.sin(th / 2)]])
    np.testing.assert_allclose(d, exact, atol=1e-13)
    du = G.derivative_matrices(G.U(0.3, 0.2, 0.1), [0])
    assert len(du) == 3


def _angled_gates(rng):
    a = lambda: float(rng.uniform(-3, 3))  # noqa: E731
    one = [G.Rx(a()), G.Ry(a()), G.Rz(a()), G.PhaseShift(a()), G.U(a(), a(), a()), G.GPi(a()), G.GPi2(a()),
           G.PRx(a(), a())]
    two = [G.MS(a(), a(), a()), G.PSwap(a()), G.XX(a()), G.YY(a()), G.XY(a()), G.ZZ(a()), G.CPhaseShift(a()),
           G.CPhaseShift00(a()), G.CPhaseShift01(a()), G.CPhaseShift10(a()), G.SingleExcitation(a()),
           G.SingleExcitationPlus(a()), G.SingleExcitationMinus(a()), G.FermionicSWAP(a())]
    out = [(g, [0]) for g in one] + [(g, [1, 0]) for g in two]
    out += [(G.MultiRZ(a()), [2, 0, 1]), (G.GPhase(2, a()), [0, 1]), (G.DoubleExcitation(a()), [0, 1, 2, 3])]
    out += [(G.Control(G.Ry(a()), (1,)), [1, 0]), (G.Control(G.U(a(), a(), a()), (0,)), [0, 2]),
            (G.Control(G.XY(a()), (1, 0)), [3, 0, 2, 1])]
    out += [(G.Rx(a(), pow_exponent=0.37), [0]), (G.CPhaseShift(a(), pow_exponent=-1.5), [0, 1])]
    return out


def test_derivative_matrices_every_angled_gate():
    """The exact derivative rules (shifted matrices) agree with an
    independent 8th-order central difference for every angle of every gate
    family, controlled forms and non-unit exponents included."""
    rng = np.random.default_rng(5)
    for g, t in _angled_gates(rng):
        d = G.derivative_matrices(g, t)
        assert len(d) == len(g.angles)
        for i in range(len(g.angles)):
            np.testing.assert_allclose(d[i], G._fd_derivative(g, t, i), atol=1e-12, err_msg=repr(g))
    sub = G.derivative_matrices(G.U(0.3, 0.2, 0.1), [0], [2])
    np.testing.assert_allclose(sub[0], G.derivative_matrices(G.U(0.3, 0.2, 0.1), [0])[2])


def test_derivative_matrices_closed_forms():
    """PhaseShift, U(theta, phi, lambda) in phi and a controlled Rx in
    closed form."""
    p = 0.61
    np.testing.assert_allclose(G.derivative_matrices(G.PhaseShift(p), [0])[0],
                               np.diag([0, 1j * np.exp(1j * p)]), atol=1e-15)
    th, ph, la = 0.4, -1.1, 2.2
    du = G.derivative_matrices(G.U(th, ph, la), [0])[1]
    exact = np.array([[0, 0], [1j * np.exp(1j * ph) * math.sin(th / 2), 1j * np.exp(1j * (ph + la)) * math.cos(th / 2)]])
    np.testing.assert_allclose(du, exact, atol=1e-15)


def test_workload_sizes():
    ins, th = W.ry_qft(20)
    assert len(ins) == 20 + 20 + 190 + 10
    bw = W.brickwall(30, 20)
    assert len(bw) == 600 + 290
    h, pm, th = W.hea(24, 8)
    assert len(pm) == 384 and len(h) == 8 * (48 + 23)
    words, c = W.tfim(24)
    assert len(words) == 23 + 24


def test_multi_qubit_pauli_channel():
    """MultiQubitPauliChannel{N} (src/noises.jl:71-90): for N = 2 the Kraus
    matrices of TwoQubitPauliChannel, trace preserving; other N raise."""
    from bsim import noises as Nz

    probs = {"XY": 0.1, "ZZ": 0.05, "IX": 0.02}
    a, b = Nz.MultiQubitPauliChannel(probs), Nz.TwoQubitPauliChannel(probs)
    assert a.n_qubits == 2
    ka = a.kraus_rep()
    for x, y in zip(ka, b.kraus_rep()):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_allclose(sum(k.conj().T @ k for k in ka), np.eye(4), atol=1e-15)
    with pytest.raises(NotImplementedError):
        Nz.MultiQubitPauliChannel({"XYZ": 0.1}).kraus_rep()


def test_fast_lowering_matches_general_path():
    """lower_instructions' shortcuts (Pauli rotations at exponent 1 and
    fixed gates with reused kernel calls) produce the op records and matrix
    bytes of the general kernel_call / derivative_matrices path, with and
    without a parameter map."""
    import numpy as np

    from bsim import _lib as L
    from bsim import gates as G
    from bsim import workloads as W
    from bsim.simulators import Instruction, lower_instructions

    from circuits import random_circuit

    rng = np.random.default_rng(3)
    ins, pmap, _ = W.hea(9, 3, seed=4)
    cases = [(ins, pmap), (ins, None), (random_circuit(7, 200, rng), None),
             ([Instruction(G.Rx(0.3, pow_exponent=2.0), [1]), Instruction(G.CNot(pow_exponent=0.5), [0, 2]),
               Instruction(G.Rz(-1.1), [3]), Instruction(G.H(), [4]), Instruction(G.CCNot(), [0, 1, 2])], None)]
    for instrs, pm in cases:
        a = lower_instructions(list(instrs), L.BSIM_SV, param_map=pm, fast=True)
        b = lower_instructions(list(instrs), L.BSIM_SV, param_map=pm, fast=False)
        a._finish()
        b._finish()
        assert a.n == b.n
        for f in ("type", "n_targets", "n_controls", "targets", "controls", "control_values", "n_params",
                  "param_index"):
            np.testing.assert_array_equal(a._rec[f], b._rec[f], err_msg=f)
        assert a._buf.tobytes() == b._buf.tobytes()


def test_bind_matches_lowering_of_rebound_instructions():
    """OpList.bind(theta) on an op list lowered with a param_map gives the
    op records and matrix values of lowering the rebound instructions
    (adjoint.with_params) from scratch: Pauli rotations (recomputed
    together), other parametrized gates (U, PhaseShift, controlled
    rotations, XX, ZZ, CPhaseShift) and gates without parameters; binding
    twice in a row and binding back to the original angles included."""
    from bsim import adjoint as A

    from circuits import one_qubit_runs

    rng = np.random.default_rng(11)
    cases = [W.hea(9, 3, seed=4), one_qubit_runs(8, 3, 5)]
    ins, pmap, p = [], {}, 0
    for q in range(6):
        for g, k in ((G.U, 3), (G.PhaseShift, 1), (G.Rx, 1)):
            ins.append(Instruction(g(*rng.uniform(-3, 3, k)), [q]))
            for j in range(k):
                pmap[(len(ins) - 1, j)] = p
                p += 1
        g = [G.XX, G.ZZ, G.CPhaseShift][q % 3](float(rng.uniform(-3, 3)))
        ins.append(Instruction(g, [q, (q + 1) % 6]))
        pmap[(len(ins) - 1, 0)] = p
        p += 1
        ins.append(Instruction(G.Control(G.Ry(float(rng.uniform(-3, 3))), (0,)), [(q + 2) % 6, q]))
        pmap[(len(ins) - 1, 0)] = p
        p += 1
        ins.append(Instruction(G.CNot(), [q, (q + 3) % 6]))
    th0 = np.zeros(p)
    for (i, j), k in pmap.items():
        th0[k] = ins[i].operator.angles[j]
    cases.append((ins, pmap, th0))
    for instrs, pm, th in cases:
        ol = lower_instructions(list(instrs), L.BSIM_SV, param_map=pm)
        ol.ptr  # the records and buffer exist before rebinding
        for t in (th + rng.uniform(-1, 1, th.size), th - 0.25, th):
            ol.bind(t)
            ref = lower_instructions(A.with_params(list(instrs), pm, t), L.BSIM_SV, param_map=pm)
            ref._finish()
            assert ol.n == ref.n
            for f in ("type", "n_targets", "n_controls", "targets", "controls", "control_values", "n_params",
                      "param_index"):
                np.testing.assert_array_equal(ol._rec[f], ref._rec[f], err_msg=f)
            np.testing.assert_array_equal(ol._buf, ref._buf)


def test_merge_width_setting():
    """bsim_set_merge_width: -1 reads, 0 and 3..6 set, anything else is refused."""
    import ctypes

    import pytest

    from bsim import _lib as L

    old = ctypes.c_int()
    L.call("bsim_set_merge_width", -1, ctypes.byref(old))
    assert old.value in (0, 3, 4, 5, 6)
    try:
        for v in (0, 3, 5, 6):
            L.call("bsim_set_merge_width", v, None)
            got = ctypes.c_int()
            L.call("bsim_set_merge_width", -1, ctypes.byref(got))
            assert got.value == v
        for bad in (-2, 1, 2, 7):
            with pytest.raises(L.BsimError):
                L.call("bsim_set_merge_width", bad, None)
    finally:
        L.call("bsim_set_merge_width", old.value, None)
