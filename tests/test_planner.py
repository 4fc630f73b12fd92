"""The fused planner on the CPU: libbsim's own host planner (lower_ops +
make_plan) is run on random circuits and its plan is interpreted op by op
(tests/native/plan_interp.cpp) on a full state vector; the result must match
the oracle.  This pins the planner's pass/sweep ordering, the 2-register
block fusion, the sign-op merging and the per-thread/per-register split of
the index tests without a GPU."""
import ctypes
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from bsim import _lib as L
from bsim import gates as G
from bsim import noises as Nz
from bsim import workloads as W
from bsim.simulators import Instruction, lower_instructions

from circuits import chain_circuit, dm_sparse_noise, grcs16, random_circuit
from conftest import GOLDEN, ROOT

HERE = Path(__file__).resolve().parent
SRC = HERE / "native" / "plan_interp.cpp"
OUT = HERE / "native" / "_build" / "libplan_interp.so"


@pytest.fixture(scope="module")
def interp():
    lib = ROOT / "braketsimulator.jl_amd" / "lib" / "libbsim.so"
    if not lib.exists():
        pytest.skip("libbsim not built")
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    csrc = ROOT / "braketsimulator.jl_amd" / "csrc"
    deps = [SRC, lib] + list(csrc.glob("*.h"))
    if not OUT.exists() or OUT.stat().st_mtime < max(p.stat().st_mtime for p in deps):
        OUT.parent.mkdir(exist_ok=True)
        cmd = [hipcc, "-O1", "-std=c++17", "-fPIC", "-shared", "-I", str(ROOT / "include"), "-I", str(csrc),
               str(SRC), "-o", str(OUT), "-L", str(lib.parent), "-lbsim", f"-Wl,-rpath,{lib.parent}"]
        subprocess.run(cmd, check=True, capture_output=True)
    L.load()
    h = ctypes.CDLL(str(OUT))
    for f in (h.planexec, h.planexec_dm):
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_int, ctypes.POINTER(L.Op), ctypes.c_int, ctypes.c_void_p]
    h.plan_singles.restype = ctypes.c_int
    h.plan_singles.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(L.Op), ctypes.c_int]
    h.plan_merged.restype = ctypes.c_int
    h.plan_merged.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(L.Op), ctypes.c_int]
    h.plan_passes.restype = ctypes.c_int
    h.plan_passes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(L.Op), ctypes.c_int]
    h.plan_chain_stats.restype = ctypes.c_int
    h.plan_chain_stats.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(L.Op), ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int)]
    h.plan_kind_count.restype = ctypes.c_int
    h.plan_kind_count.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(L.Op), ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int]
    h.adjexec.restype = ctypes.c_int
    h.adjexec.argtypes = [ctypes.c_int, ctypes.POINTER(L.Op), ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                          ctypes.c_void_p, ctypes.c_void_p]
    h.plan_superop_forms.restype = ctypes.c_int
    h.plan_superop_forms.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(L.Op), ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int)]
    h.plan_adj_stats.restype = ctypes.c_int
    h.plan_adj_stats.argtypes = [ctypes.c_int, ctypes.POINTER(L.Op), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    return h


def _run(h, n, ins):
    ol = lower_instructions(ins, L.BSIM_SV)
    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi[0] = 1
    rc = h.planexec(n, ol.ptr, ol.n, psi.ctypes.data)
    assert rc == 0, f"planexec returned {rc} (-10/-11: an index-test split disagrees with the direct test)"
    return psi


@pytest.fixture
def tile_config():
    t, lo, rb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L.load()
    L.call("bsim_get_tile_config", ctypes.byref(t), ctypes.byref(lo))
    L.call("bsim_set_sweep_bits", 0, ctypes.byref(rb))
    yield
    L.call("bsim_set_tile_config", t.value, lo.value)
    L.call("bsim_set_sweep_bits", rb.value, None)


KINDS = ("1q", "2q", "ctrl", "diag", "pow")


@pytest.mark.parametrize("seed", range(6))
def test_planner_random_circuits(interp, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(4, 14))
    ins = random_circuit(n, 80, rng, kinds=KINDS)
    assert np.max(np.abs(_run(interp, n, ins) - O.evolve_sv(n, ins))) < 1e-12


@pytest.mark.parametrize("bits", [3, 4])
@pytest.mark.parametrize("tile,low", [(6, 0), (8, 3), (12, 4), (12, 10)])
def test_planner_tile_configs(interp, tile_config, tile, low, bits):
    L.call("bsim_set_tile_config", tile, low)
    L.call("bsim_set_sweep_bits", bits, None)
    rng = np.random.default_rng(tile * 31 + low)
    n = 13
    ins = random_circuit(n, 70, rng, kinds=KINDS)
    assert np.max(np.abs(_run(interp, n, ins) - O.evolve_sv(n, ins))) < 1e-12


def test_planner_brickwall_and_grcs(interp):
    ins = W.brickwall(14, 12, seed=4)
    assert np.max(np.abs(_run(interp, 14, ins) - O.evolve_sv(14, ins))) < 1e-12
    g = grcs16(GOLDEN / "grcs_16.qasm")
    assert np.max(np.abs(_run(interp, 16, g) - O.evolve_sv(16, g))) < 1e-12


def test_planner_cz_with_global_bits(interp):
    """Minimal case of a planner bug found on the GPU: a CZ whose second bit
    lies outside the tile must flush the 2-register block of its tile bit."""
    h = np.pi / 2
    ins = [Instruction(G.Rx(h), [2]), Instruction(G.Rx(h), [8]), Instruction(G.Rx(h), [13]),
           Instruction(G.CZ(), [8, 4]), Instruction(G.Ry(h), [1]), Instruction(G.Rx(h), [3])]
    assert np.max(np.abs(_run(interp, 14, ins) - O.evolve_sv(14, ins))) < 1e-12


def test_planner_sign_terms(interp):
    """CZ layers at several bit distances merge into one multi-term sign op;
    every split sign table is checked against the direct parity inside the
    interpreter."""
    n = 12
    ins = [Instruction(G.H(), [q]) for q in range(n)]
    for d in (1, 2, 3, 5, 7):
        ins += [Instruction(G.CZ(), [q, q + d]) for q in range(0, n - d, 2)]
    ins += [Instruction(G.Rx(0.3 * q), [q]) for q in range(n)]
    assert np.max(np.abs(_run(interp, n, ins) - O.evolve_sv(n, ins))) < 1e-12


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("tile,low,bits", [(12, 3, 4), (6, 0, 3), (8, 2, 4)])
def test_planner_register_diagonals(interp, tile_config, seed, tile, low, bits):
    """Diagonal gates whose bits are all register bits of their sweep run as
    register-diagonal tile ops (one factor per register); controls on
    register and on thread bits split like those of dense ops."""
    L.call("bsim_set_tile_config", tile, low)
    L.call("bsim_set_sweep_bits", bits, None)
    rng = np.random.default_rng(500 + seed)
    n = 8
    ins = [Instruction(G.H(), [q]) for q in range(n)]
    for q in range(n - 1):
        a = rng.uniform(-3, 3, 4)
        ins.append(Instruction(G.Rx(a[0]), [q]))
        ins.append(Instruction(G.Control(G.Rz(a[1]), (int(rng.integers(2)),)), [q + 1, q]))
        ins.append(Instruction(G.Control(G.PhaseShift(a[2]), (int(rng.integers(2)),)), [(q + 4) % n, q]))
        ins.append(Instruction(G.Rz(a[3]), [q]))
        ins.append(Instruction(G.CPhaseShift(a[0] - a[1]), [q, q + 1]))
        ins.append(Instruction(G.Ry(a[2]), [q + 1]))
    ol = lower_instructions(ins, L.BSIM_SV)
    assert interp.plan_kind_count(n, L.BSIM_SV, ol.ptr, ol.n, 6, 0) > 0
    assert interp.plan_kind_count(n, L.BSIM_SV, ol.ptr, ol.n, 6, 1) > 0
    np.testing.assert_allclose(_run(interp, n, ins), O.evolve_sv(n, ins), atol=1e-12)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("tile,low", [(12, 3), (6, 0), (8, 2)])
def test_planner_x_swaps(interp, tile_config, seed, tile, low):
    """In 3-register sweeps X, CNOT, Toffoli and 0-controlled X run as
    register swaps (with the control split of dense ops) between rotations."""
    bits = 3
    L.call("bsim_set_tile_config", tile, low)
    L.call("bsim_set_sweep_bits", bits, None)
    rng = np.random.default_rng(700 + seed)
    n = 9
    ins = [Instruction(G.Ry(float(rng.uniform(-3, 3))), [q]) for q in range(n)]
    for _ in range(30):
        q = [int(x) for x in rng.choice(n, 3, replace=False)]
        r = rng.random()
        if r < 0.25:
            ins.append(Instruction(G.X(), [q[0]]))
        elif r < 0.55:
            ins.append(Instruction(G.CNot(), q[:2]))
        elif r < 0.7:
            ins.append(Instruction(G.CCNot(), q))
        elif r < 0.8:
            ins.append(Instruction(G.Control(G.X(), (0,)), q[:2]))
        else:
            ins.append(Instruction(G.Rx(float(rng.uniform(-3, 3))), [q[0]]))
        ins.append(Instruction(G.Rz(float(rng.uniform(-3, 3))), [q[1]]))
    ol = lower_instructions(ins, L.BSIM_SV)
    assert interp.plan_kind_count(n, L.BSIM_SV, ol.ptr, ol.n, 7, 0) > 0
    assert interp.plan_kind_count(n, L.BSIM_SV, ol.ptr, ol.n, 7, 1) > 0
    np.testing.assert_allclose(_run(interp, n, ins), O.evolve_sv(n, ins), atol=1e-12)


def _run_dm(h, n, ins, mode="reference"):
    ol = lower_instructions(ins, L.BSIM_DM, mode)
    assert h.plan_singles(n, L.BSIM_DM, ol.ptr, ol.n) == 0, "every op should be planned into tile passes"
    rho = np.zeros(4 ** n, dtype=np.complex128)
    rho[0] = 1
    rc = h.planexec_dm(n, ol.ptr, ol.n, rho.ctypes.data)
    assert rc == 0, f"planexec_dm returned {rc} (-12: a Kraus block-test split disagrees with the direct test)"
    return rho


def _noisy_circuit(n, m, rng):
    one = [lambda: Nz.Depolarizing(float(rng.uniform(0, 0.3))), lambda: Nz.AmplitudeDamping(float(rng.uniform(0, 0.5))),
           lambda: Nz.PhaseDamping(float(rng.uniform(0, 0.5))), lambda: Nz.BitFlip(float(rng.uniform(0, 0.5))),
           lambda: Nz.GeneralizedAmplitudeDamping(float(rng.uniform(0, 0.4)), float(rng.uniform(0, 1)))]
    two = [lambda: Nz.TwoQubitDepolarizing(float(rng.uniform(0, 0.3))), lambda: Nz.TwoQubitDephasing(float(rng.uniform(0, 0.3)))]
    ins = random_circuit(n, m, rng, kinds=("1q", "2q", "diag"))
    out = []
    for g in ins:
        out.append(g)
        r = rng.random()
        if r < 0.3:
            out.append(Instruction(one[rng.integers(len(one))](), [int(rng.integers(n))]))
        elif r < 0.5:
            t = [int(x) for x in rng.choice(n, 2, replace=False)]
            out.append(Instruction(two[rng.integers(len(two))](), t))
        elif r < 0.55:
            out.append(Instruction(Nz.PhaseFlip(float(rng.uniform(0, 0.5))), [int(rng.integers(n))]))
    return out


@pytest.mark.parametrize("mode", ["reference", "physical"])
@pytest.mark.parametrize("seed", range(5))
def test_planner_dm_kraus_in_tiles(interp, seed, mode):
    """1- and 2-qubit Kraus channels planned into the tile passes:
    reference semantics as predicated superoperators, physical 2-qubit
    channels as 16x16 tile ops.  The interpreted plan must match the
    oracle's Kraus loops."""
    rng = np.random.default_rng(300 + seed)
    n = int(rng.integers(3, 7))
    ins = _noisy_circuit(n, 30, rng)
    np.testing.assert_allclose(_run_dm(interp, n, ins, mode), O.evolve_dm(n, ins, mode), atol=1e-12)


@pytest.mark.parametrize("mode", ["reference", "physical"])
@pytest.mark.parametrize("tile,low", [(6, 0), (8, 3), (12, 4)])
def test_planner_dm_noisy_brickwall(interp, tile_config, tile, low, mode):
    L.call("bsim_set_tile_config", tile, low)
    ins = W.brickwall(6, 3, seed=2, noise=True)
    np.testing.assert_allclose(_run_dm(interp, 6, ins, mode), O.evolve_dm(6, ins, mode), atol=1e-12)


def _generic_channel(k, rng):
    """Two Kraus matrices from a random isometry: a channel whose block
    superoperator links row == column states of its qubits with the rest,
    so the planner must keep it a full barrier."""
    d = 2 ** k
    q = np.linalg.qr(rng.normal(size=(2 * d, d)) + 1j * rng.normal(size=(2 * d, d)))[0]
    return Nz.Kraus([q[:d], q[d:]])


@pytest.mark.parametrize("mode", ["reference", "physical"])
@pytest.mark.parametrize("tile,low", [(6, 0), (8, 3)])
@pytest.mark.parametrize("seed", range(4))
def test_planner_dm_kraus_reordering(interp, tile_config, tile, low, mode, seed):
    """Reference-semantics Kraus ops move past each other and past diagonal
    ops on other qubits when both keep the row/column pair predicates
    (depolarizing, damping, dephasing, Pauli channels), and stay barriers
    when they do not (generic channels).  Small tiles force many passes, so
    ops are taken ahead of held-back ones; the interpreted plan must match
    the oracle's loops in program order."""
    L.call("bsim_set_tile_config", tile, low)
    rng = np.random.default_rng(900 + seed)
    n = 6
    ins = []
    for _ in range(3):
        ins += random_circuit(n, 8, rng, kinds=("1q", "2q", "diag"))
        for q in rng.permutation(n):
            r = rng.random()
            if r < 0.25:
                ins.append(Instruction(_generic_channel(1, rng), [int(q)]))
            elif r < 0.6:
                ins.append(Instruction(Nz.AmplitudeDamping(float(rng.uniform(0, 0.4))), [int(q)]))
            else:
                ins.append(Instruction(Nz.Depolarizing(float(rng.uniform(0, 0.3))), [int(q)]))
        for q in range(seed % 2, n - 1, 2):
            ins.append(Instruction(G.CZ(), [q, q + 1]))
            ins.append(Instruction(_generic_channel(2, rng) if rng.random() < 0.3 else
                                   Nz.TwoQubitDepolarizing(float(rng.uniform(0, 0.2))), [q, q + 1]))
        ins.append(Instruction(Nz.PhaseFlip(0.1), [int(rng.integers(n))]))
    np.testing.assert_allclose(_run_dm(interp, n, ins, mode), O.evolve_dm(n, ins, mode), atol=1e-12)


def test_planner_dm_kraus_pass_count(interp):
    """The 14-qubit noisy brick-wall of depth 10 in reference semantics: with
    the channels commuting past each other the plan needs 51 passes (90
    when every Kraus op is a barrier); the same circuit with generic
    channels in place of AmplitudeDamping keeps them barriers."""
    ins = W.brickwall(14, 10, seed=0, noise=True)
    ol = lower_instructions(ins, L.BSIM_DM, "reference")
    good = interp.plan_passes(14, L.BSIM_DM, ol.ptr, ol.n)
    assert 0 < good <= 52
    rng = np.random.default_rng(5)
    gen = [Instruction(_generic_channel(1, rng), g.target) if isinstance(g.operator, Nz.AmplitudeDamping) else g
           for g in ins]
    ol = lower_instructions(gen, L.BSIM_DM, "reference")
    assert interp.plan_passes(14, L.BSIM_DM, ol.ptr, ol.n) > good + 10


# ---- fused adjoint: the backward plan ---------------------------------------

_PAULI = {"X": np.array([[0, 1], [1, 0]], dtype=complex), "Y": np.array([[0, -1j], [1j, 0]]),
          "Z": np.array([[1, 0], [0, -1]], dtype=complex)}


def _h_apply(psi, n, words, coeffs):
    out = np.zeros_like(psi)
    for w, c in zip(words, coeffs):
        v = psi.copy()
        for q, ch in enumerate(w):
            if ch != "I":
                v = O.np_apply(v, n, _PAULI[ch], [q])
        out += c * v
    return out


def _energy(n, ins, words, coeffs):
    psi = O.evolve_sv(n, ins)
    return sum(c * O.sv_expect_pauli(psi, n, w) for w, c in zip(words, coeffs))


def _adj_interp(h, n, ins, pmap, words, coeffs, n_params):
    """Gradient from the interpreted backward plan (libbsim's own
    adjoint_back_plan: reversed U^dagger ops, payloads of dU)."""
    psi = O.evolve_sv(n, ins).astype(np.complex128)
    lam = _h_apply(psi, n, words, coeffs)
    e_h = float(np.vdot(psi, lam).real)
    assert abs(e_h - _energy(n, ins, words, coeffs)) < 1e-12  # pins _h_apply against the oracle's <P>
    ol = lower_instructions(list(ins), L.BSIM_SV, param_map=pmap)
    grad = np.zeros(n_params)
    rc = h.adjexec(n, ol.ptr, ol.n, n_params, psi.ctypes.data, lam.ctypes.data, grad.ctypes.data)
    assert rc == 0, f"adjexec returned {rc}"
    return grad


@pytest.mark.parametrize("n,layers", [(6, 2), (13, 2)])
def test_adjoint_back_plan_hea(interp, n, layers):
    """The backward plan of the hardware-efficient ansatz, interpreted on the
    CPU, vs the oracle's parameter shift (exact for RY/RZ).  13 qubits: two
    tiles of 11 bits per pass, parameters on bits outside the first tile."""
    from bsim import adjoint as A

    ins, pmap, th = W.hea(n, layers, seed=n)
    words, coeffs = W.tfim(n)
    grad = _adj_interp(interp, n, ins, pmap, words, coeffs, len(th))
    ps = np.zeros(len(th))
    for p in range(len(th)):
        tp = th.copy(); tp[p] += np.pi / 2
        tm = th.copy(); tm[p] -= np.pi / 2
        ps[p] = 0.5 * (_energy(n, A.with_params(ins, pmap, tp), words, coeffs)
                       - _energy(n, A.with_params(ins, pmap, tm), words, coeffs))
    np.testing.assert_allclose(grad, ps, atol=1e-10)


@pytest.mark.parametrize("n", [7, 12])
def test_adjoint_back_plan_general_gates(interp, n):
    """U(theta, phi, lambda) (three products per op), ZZ / CPhaseShift
    (diagonal), controlled Rx with either control value, XX (dense 2-qubit),
    Rx(0) (the identity, whose derivative is not diagonal) and a
    parameter-free 3-qubit unitary (a non-tile step), vs central differences
    of the oracle."""
    from bsim import adjoint as A
    from circuits import random_unitary

    rng = np.random.default_rng(40 + n)
    ins, pmap = [], {}
    p = 0
    for layer in range(3):
        for q in range(n):
            ins.append(Instruction(G.U(*rng.uniform(-3, 3, 3)), [q]))
            for j in range(3):
                pmap[(len(ins) - 1, j)] = p
                p += 1
        ins.append(Instruction(G.Rx(0.0), [layer % n]))
        pmap[(len(ins) - 1, 0)] = p
        p += 1
        for q in range(n - 1):
            a = float(rng.uniform(-3, 3))
            g = [G.ZZ(a), G.CPhaseShift(a), G.Control(G.Rx(a), (1,)), G.XX(a), G.Control(G.Ry(a), (0,))][q % 5]
            t = [q, q + 1] if q % 2 == 0 else [q + 1, (q + 5) % n]
            if t[0] == t[1]:
                t = [q, q + 1]
            ins.append(Instruction(g, t))
            pmap[(len(ins) - 1, 0)] = p
            p += 1
        ins.append(Instruction(G.Unitary(random_unitary(3, rng)), [0, n // 2, n - 1]))
    th = np.zeros(p)
    for (i, j), k in pmap.items():
        th[k] = ins[i].operator.angles[j]
    words, coeffs = W.tfim(n)
    words = list(words) + ["Y" + "I" * (n - 2) + "X"]
    coeffs = list(coeffs) + [0.3]
    grad = _adj_interp(interp, n, ins, pmap, words, coeffs, p)
    hstep = 1e-5
    fd = np.zeros(p)
    for k in range(p):
        tp = th.copy(); tp[k] += hstep
        tm = th.copy(); tm[k] -= hstep
        fd[k] = (_energy(n, A.with_params(ins, pmap, tp), words, coeffs)
                 - _energy(n, A.with_params(ins, pmap, tm), words, coeffs)) / (2 * hstep)
    np.testing.assert_allclose(grad, fd, atol=1e-8)


def _adj_stats(h, n, ins, pmap):
    ol = lower_instructions(list(ins), L.BSIM_SV, param_map=pmap)
    out = (ctypes.c_int * 5)()
    assert h.plan_adj_stats(n, ol.ptr, ol.n, out) == 0
    return list(out)


@pytest.mark.parametrize("n,layers,seed", [(7, 2, 0), (12, 3, 1), (13, 2, 2)])
def test_adjoint_merged_one_qubit_runs(interp, n, layers, seed):
    """Runs of one-qubit ops on one qubit merge into one backward op, their
    products taken before or after it (Pauli-coded) or after it (payloads):
    every run shape of circuits.one_qubit_runs, interpreted on the CPU
    (coded products cross-checked against their payloads inside adjexec),
    vs central differences of the oracle; and the merges happen."""
    from bsim import adjoint as A
    from circuits import one_qubit_runs

    ins, pmap, th = one_qubit_runs(n, layers, seed)
    words, coeffs = W.tfim(n)
    words = list(words) + ["Y" + "I" * (n - 2) + "X"]
    coeffs = list(coeffs) + [0.3]
    st = _adj_stats(interp, n, ins, pmap)
    assert st[1] >= n * layers // 2, st   # merged runs
    assert st[3] > 0 and st[4] > 0, st    # coded products on both sides
    grad = _adj_interp(interp, n, ins, pmap, words, coeffs, len(th))
    hstep = 1e-5
    fd = np.zeros(len(th))
    for k in range(len(th)):
        tp = th.copy(); tp[k] += hstep
        tm = th.copy(); tm[k] -= hstep
        fd[k] = (_energy(n, A.with_params(ins, pmap, tp), words, coeffs)
                 - _energy(n, A.with_params(ins, pmap, tm), words, coeffs)) / (2 * hstep)
    np.testing.assert_allclose(grad, fd, atol=1e-8)


def test_adjoint_merge_hea_counts(interp):
    """The hardware-efficient ansatz's RZ^dagger RY^dagger pairs: every pair
    one merged op with its Z product before and its Y product after."""
    n, layers = 10, 3
    ins, pmap, th = W.hea(n, layers, seed=3)
    st = _adj_stats(interp, n, ins, pmap)
    assert st[1] == n * layers and st[2] == n * layers, st
    assert st[3] == n * layers and st[4] == n * layers, st
    assert st[0] == len(ins) - n * layers, st


@pytest.mark.parametrize("seed", range(3))
def test_planner_merges_dense_runs(interp, seed):
    """Runs of adjacent dense ops around 3+-target gates are multiplied into
    one op of at most 6 bits (a matrix-core launch instead of several HBM
    passes; vectors of 20+ bits): the interpreted plan, merged steps
    included, matches the oracle on random circuits with 3- and 4-target
    unitaries, and merging happens; below 20 bits nothing is merged."""
    from circuits import random_unitary

    rng = np.random.default_rng(700 + seed)
    n = 20
    ins = []
    for _ in range(5):
        k = int(rng.integers(3, 5))
        q = [int(x) for x in rng.choice(n, k, replace=False)]
        ins.append(Instruction(G.Unitary(random_unitary(k, rng)), q))
        # neighbours: inside the block, another 3-target gate sharing a qubit, a far gate
        ins.append(Instruction(G.Rx(float(rng.uniform(-3, 3))), [q[0]]))
        if rng.random() < 0.6:
            q2 = [q[1]] + [int(x) for x in rng.choice([x for x in range(n) if x not in q], 2, replace=False)]
            ins.append(Instruction(G.Unitary(random_unitary(3, rng)), q2))
        ins.append(Instruction(G.CZ(), [int(x) for x in rng.choice(n, 2, replace=False)]))
    ol = lower_instructions(ins, L.BSIM_SV)
    assert interp.plan_merged(n, L.BSIM_SV, ol.ptr, ol.n) > 0
    assert np.max(np.abs(_run(interp, n, ins) - O.evolve_sv(n, ins))) < 1e-12
    small = lower_instructions([Instruction(i.operator, [q % 9 for q in i.target]) for i in ins
                                if len(set(q % 9 for q in i.target)) == len(i.target)], L.BSIM_SV)
    assert interp.plan_merged(9, L.BSIM_SV, small.ptr, small.n) == 0


def test_planner_merges_dm_gate_halves(interp):
    """A 3-target gate on an 11-qubit density matrix is U on 3 column bits
    and conj(U) on 3 row bits: one merged 6-bit op (22-bit vector); the
    interpreted plan matches the oracle."""
    from circuits import random_unitary

    rng = np.random.default_rng(5)
    n = 11
    ins = [Instruction(G.H(), [q]) for q in range(n)]
    ins += [Instruction(G.Unitary(random_unitary(3, rng)), [0, 3, 4]), Instruction(G.Ry(0.7), [2]),
            Instruction(G.Unitary(random_unitary(3, rng)), [4, 1, 2])]
    ol = lower_instructions(ins, L.BSIM_DM)
    assert interp.plan_merged(n, L.BSIM_DM, ol.ptr, ol.n) >= 2
    rho = np.zeros(4 ** n, dtype=np.complex128)
    rho[0] = 1
    rc = interp.planexec_dm(n, ol.ptr, ol.n, rho.ctypes.data)
    assert rc == 0
    np.testing.assert_allclose(rho, O.evolve_dm(n, ins), atol=1e-12)


# ---- register chains ---------------------------------------------------------

def _chain_stats(h, n, ins, kind=L.BSIM_SV, mode="reference"):
    ol = lower_instructions(ins, kind, noise_mode=mode)
    out = (ctypes.c_int * 7)()
    assert h.plan_chain_stats(n, kind, ol.ptr, ol.n, out) == 0
    return {"start": out[0], "mul": out[1], "apply": out[2], "psign": out[3], "chain": out[4], "split_sign": out[5],
            "longest": out[6]}


def test_planner_chain_forms(interp):
    """Every register-chain form, interpreted with its documented
    per-sub-cube meaning, against the oracle: chains of up to four
    matrices in one op, longer chains (start / multiply / per-thread sign /
    apply), and the sign ops a rewrite leaves at the end of a sweep."""
    ins = chain_circuit(13, 9, 0, 2)
    s = _chain_stats(interp, 13, ins)
    assert s["start"] and s["mul"] and s["apply"] and s["psign"], s
    assert s["chain"] and s["split_sign"] and s["longest"] == 4, s
    assert np.max(np.abs(_run(interp, 13, ins) - O.evolve_sv(13, ins))) < 1e-12
    ins = W.brickwall(14, 12, seed=4)
    s = _chain_stats(interp, 14, ins)
    assert s["chain"] >= 10 and s["start"] == 0, s
    assert np.max(np.abs(_run(interp, 14, ins) - O.evolve_sv(14, ins))) < 1e-12


def test_planner_chain_random(interp):
    """Random chain circuits (with or without chains, depending on the
    sweeps the planner forms) against the oracle."""
    chains = 0
    for seed in range(8):
        n = 12 + seed % 2
        ins = chain_circuit(n, 6 + seed, 10 + seed, 1 + seed % 3)
        s = _chain_stats(interp, n, ins)
        chains += s["chain"] + s["start"]
        assert np.max(np.abs(_run(interp, n, ins) - O.evolve_sv(n, ins))) < 1e-12, seed
    assert chains > 0


@pytest.mark.parametrize("mode", ["reference", "physical"])
def test_planner_chains_density_matrix(interp, mode):
    """Density matrices of 6 and 7 qubits (12 and 14 vector bits): chains in
    the noise-free passes, the Kraus passes around them unchanged."""
    ins = W.brickwall(6, 8, seed=6)
    assert _chain_stats(interp, 6, ins, L.BSIM_DM, mode)["chain"] > 0
    np.testing.assert_allclose(_run_dm(interp, 6, ins, mode), O.evolve_dm(6, ins, mode), atol=1e-12)
    ins = dm_sparse_noise(7, 6, 7)
    assert _chain_stats(interp, 7, ins, L.BSIM_DM, mode)["chain"] > 0
    np.testing.assert_allclose(_run_dm(interp, 7, ins, mode), O.evolve_dm(7, ins, mode), atol=1e-12)


def _cnot_circuit(n, layers, seed, cval0=False):
    """Random one-qubit gates and CNOT ladders, controls on low and high
    bits (with `cval0` one CNOT with control value 0, which keeps the
    original op list)."""
    rng = np.random.default_rng(seed)
    ins = []
    for layer in range(layers):
        for q in range(n):
            ins.append(Instruction(G.U(*rng.uniform(-3, 3, 3)), [q]))
        for q in range(layer % 2, n - 1, 2):
            ins.append(Instruction(G.CNot(), [q, q + 1] if (q + layer) % 3 else [q + 1, q]))
        ins.append(Instruction(G.CNot(), [0, n - 1]))
        ins.append(Instruction(G.CZ(), [1, n - 2]))
    if cval0:
        ins.append(Instruction(G.Control(G.X(), (0,)), [2, 5]))
    return ins


@pytest.mark.parametrize("n,layers,seed", [(12, 4, 0), (13, 5, 1), (15, 3, 2)])
def test_planner_cnot_rewrite(interp, n, layers, seed):
    """CNOT(c, t) planned as H_t CZ(c, t) H_t when every other op is free of
    controls (12-bit, 4-register passes with chains): the plan holds no
    controlled tile op and forms chains, and the interpreted plan matches
    the oracle; one CNOT with control value 0 keeps the original list."""
    ins = _cnot_circuit(n, layers, seed)
    ol = lower_instructions(ins, L.BSIM_SV)
    assert interp.plan_kind_count(n, L.BSIM_SV, ol.ptr, ol.n, 0, 1) == 0
    assert _chain_stats(interp, n, ins)["chain"] > 0
    assert np.max(np.abs(_run(interp, n, ins) - O.evolve_sv(n, ins))) < 1e-12
    hea, _, _ = W.hea(n, layers, seed=seed)
    ol = lower_instructions(hea, L.BSIM_SV)
    assert interp.plan_kind_count(n, L.BSIM_SV, ol.ptr, ol.n, 0, 1) == 0
    assert np.max(np.abs(_run(interp, n, hea) - O.evolve_sv(n, hea))) < 1e-12
    ins = _cnot_circuit(n, layers, seed, cval0=True)
    ol = lower_instructions(ins, L.BSIM_SV)
    assert interp.plan_kind_count(n, L.BSIM_SV, ol.ptr, ol.n, 0, 1) > 0
    assert np.max(np.abs(_run(interp, n, ins) - O.evolve_sv(n, ins))) < 1e-12


def test_planner_cnot_rewrite_density_matrix(interp):
    """The rewrite on a noise-free density matrix (both halves of every
    gate) against the oracle."""
    ins = _cnot_circuit(6, 4, 9)
    ol = lower_instructions(ins, L.BSIM_DM)
    assert interp.plan_kind_count(6, L.BSIM_DM, ol.ptr, ol.n, 0, 1) == 0
    np.testing.assert_allclose(_run_dm(interp, 6, ins), O.evolve_dm(6, ins, "reference"), atol=1e-12)
    # with reference Kraus channels between the gates (a CNOT list the lean
    # Kraus passes take)
    from bsim import noises as Nz

    noisy = []
    for k, i in enumerate(ins):
        noisy.append(i)
        if k % 3 == 0:
            noisy.append(Instruction(Nz.Depolarizing(0.05), [int(i.target[-1])]))
    ol = lower_instructions(noisy, L.BSIM_DM)
    assert interp.plan_kind_count(6, L.BSIM_DM, ol.ptr, ol.n, 0, 1) == 0
    for mode in ("reference", "physical"):
        np.testing.assert_allclose(_run_dm(interp, 6, noisy, mode), O.evolve_dm(6, noisy, mode), atol=1e-12)


@pytest.mark.parametrize("n,depth,seed", [(6, 6, 0), (7, 5, 1)])
def test_planner_superop_orbit_form(interp, n, depth, seed):
    """Physical two-qubit channels whose 16x16 superoperator couples only
    registers that differ by one of {0, P1, P2, P1 ^ P2} (depolarizing: a
    qubit's row and column bit flip together) plan as four 4x4 orbit blocks;
    a channel without that form (a random 2-qubit unitary as its one Kraus
    operator: U (x) conj(U) is dense) keeps the 16 columns; the interpreted
    plans match the oracle."""
    from bsim import noises as Nz
    from circuits import random_unitary

    ins = W.brickwall(n, depth, seed=seed, noise=True)
    ol = lower_instructions(ins, L.BSIM_DM, noise_mode="physical")
    out = (ctypes.c_int * 2)()
    assert interp.plan_superop_forms(n, L.BSIM_DM, ol.ptr, ol.n, out) == 0
    assert out[0] > 0 and out[1] == 0, list(out)
    np.testing.assert_allclose(_run_dm(interp, n, ins, "physical"), O.evolve_dm(n, ins, "physical"), atol=1e-12)
    rng = np.random.default_rng(seed)
    ks = [random_unitary(2, rng)]
    ins2 = list(ins[: len(ins) // 2]) + [Instruction(Nz.Kraus(ks), [1, 3])] + list(ins[len(ins) // 2:])
    ol = lower_instructions(ins2, L.BSIM_DM, noise_mode="physical")
    assert interp.plan_superop_forms(n, L.BSIM_DM, ol.ptr, ol.n, out) == 0
    assert out[0] > 0 and out[1] == 1, list(out)
    np.testing.assert_allclose(_run_dm(interp, n, ins2, "physical"), O.evolve_dm(n, ins2, "physical"), atol=1e-12)
