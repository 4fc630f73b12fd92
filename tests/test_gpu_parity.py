"""GPU parity: libbsim (HIP, gfx950) against the oracle on the same inputs.

Bar: amplitudes, probabilities and expectations within 1e-10 absolute of
the oracle; alias tables and shot indices bit-exact given the same
probabilities, wsum and Philox draws.
"""
import ctypes
import math
import sys

import numpy as np
import pytest

import oracle as O
import bsim
from bsim import _lib as L
from bsim import gates as G
from bsim import noises as Nz
from bsim import workloads as W
from bsim.observables import H as HObs, HermitianObservable, TensorProduct, X as XO, Y as YO, Z as ZO, I as IO
from bsim.simulators import DensityMatrixSimulator, Instruction, StateVectorSimulator

from circuits import fixture_ops, grcs16, random_circuit, random_noise, random_unitary
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

TOL = 1e-10
GOLD_ATOL = 6e-8


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cx(rows):
    return np.array([complex(a, b) for a, b in rows])


@pytest.mark.parametrize("mode", ["fused", "per_op", "apply_gate"])
def test_sv_golden(fixtures, mode):
    for case in fixtures["sv_simple"]:
        sim = StateVectorSimulator(case["n"], 0)
        ops = fixture_ops(case)
        if mode == "apply_gate":
            for i in ops:
                sim.apply_gate(i.operator, *i.target)
        else:
            sim.evolve(ops, mode=0 if mode == "fused" else 1)
        np.testing.assert_allclose(sim.state_vector(), _cx(case["state"]), atol=GOLD_ATOL, err_msg=case["cite"])
        np.testing.assert_allclose(sim.probabilities(), case["probs"], atol=GOLD_ATOL, err_msg=case["cite"])


@pytest.mark.parametrize("mode", [0, 1])
def test_dm_golden(fixtures, mode):
    for case in fixtures["dm_simple"]:
        n = case["n"]
        N = 2 ** n
        sim = DensityMatrixSimulator(n, 0)
        sim.evolve(fixture_ops(case), mode=mode)
        if "dm" in case:
            exp = np.array(case["dm"], dtype=np.complex128)
        else:
            exp = np.zeros((N, N), dtype=np.complex128)
            exp[case["dm_diag_one"], case["dm_diag_one"]] = 1
        np.testing.assert_allclose(sim.density_matrix(), exp, atol=GOLD_ATOL, err_msg=case["cite"])


def test_grcs16(fixtures):
    sim = StateVectorSimulator(16, 0)
    sim.evolve(grcs16(GOLDEN / "grcs_16.qasm"))
    p = sim.probabilities()
    assert abs(p[0] - fixtures["grcs16"]["first_prob"]) <= fixtures["grcs16"]["atol"]


@pytest.mark.parametrize("n,seed", [(3, 0), (6, 1), (9, 2), (12, 3), (14, 4), (17, 5)])
@pytest.mark.parametrize("mode", [0, 1])
def test_sv_random_vs_oracle(n, seed, mode):
    _say(f"sv random n={n} seed={seed} mode={mode}")
    rng = np.random.default_rng(seed)
    ins = random_circuit(n, 120, rng)
    exp = O.evolve_sv(n, ins)
    sim = StateVectorSimulator(n, 0)
    sim.evolve(ins, mode=mode)
    got = sim.state_vector()
    assert np.max(np.abs(got - exp)) < TOL


@pytest.mark.parametrize("bits", [3, 4])
@pytest.mark.parametrize("tile,low", [(6, 0), (8, 2), (10, 3), (12, 4), (13, 5), (12, 10)])
def test_tile_configs(tile, low, bits):
    _say(f"tile config {tile} {low} sweep bits {bits}")
    old_bits = ctypes.c_int()
    L.call("bsim_set_sweep_bits", bits, ctypes.byref(old_bits))
    n = 16
    rng = np.random.default_rng(tile * 7 + low)
    ins = random_circuit(n, 150, rng)
    exp = O.evolve_sv(n, ins)
    t0, lo0 = ctypes.c_int(), ctypes.c_int()
    L.call("bsim_get_tile_config", ctypes.byref(t0), ctypes.byref(lo0))
    L.call("bsim_set_tile_config", tile, low)
    try:
        sim = StateVectorSimulator(n, 0)
        sim.evolve(ins)
        got = sim.state_vector()
    finally:
        L.call("bsim_set_tile_config", t0.value, lo0.value)
        L.call("bsim_set_sweep_bits", old_bits.value, None)
    assert np.max(np.abs(got - exp)) < TOL


def test_per_gate_api_vs_oracle():
    n = 13
    rng = np.random.default_rng(9)
    ins = random_circuit(n, 80, rng)
    exp = O.evolve_sv(n, ins)
    sim = StateVectorSimulator(n, 0)
    for i in ins:
        sim.apply_gate(i.operator, *i.target)
    assert np.max(np.abs(sim.state_vector() - exp)) < TOL


@pytest.mark.parametrize("noise_mode", ["reference", "physical"])
@pytest.mark.parametrize("n,seed", [(2, 0), (4, 1), (7, 2)])
@pytest.mark.parametrize("mode", [0, 1])
def test_dm_random_vs_oracle(noise_mode, n, seed, mode):
    _say(f"dm random n={n} seed={seed} {noise_mode} mode={mode}")
    rng = np.random.default_rng(50 + seed)
    ops = []
    for _ in range(30):
        ops += random_circuit(n, 2, rng)
        ops.append(random_noise(n, rng))
    exp = O.evolve_dm(n, ops, mode=noise_mode)
    sim = DensityMatrixSimulator(n, 0, noise_mode=noise_mode)
    sim.evolve(ops, mode=mode)
    got = sim.density_matrix().ravel(order="F")
    assert np.max(np.abs(got - exp)) < TOL
    np.testing.assert_allclose(sim.probabilities(), O.dm_probabilities(exp, n), atol=TOL)


@pytest.mark.parametrize("n,depth", [(6, 4), (9, 3)])
def test_dm_reference_kraus_fused_into_tiles(n, depth):
    """Reference-semantics 1- and 2-qubit Kraus channels run inside the
    fused tile passes (no separate Kraus launches) and still reproduce the
    reference's block-diagonal update."""
    _say(f"dm noisy brickwall n={n} depth={depth} reference Kraus in tiles")
    ins = W.brickwall(n, depth, seed=n, noise=True)
    ol = bsim.lower_instructions(ins, L.BSIM_DM, "reference")
    st = (ctypes.c_int64 * 4)()
    L.call("bsim_plan_stats", 2 * n, 1, ol.ptr, ol.n, st)
    assert st[1] == 0
    exp = O.evolve_dm(n, ins, mode="reference")
    sim = DensityMatrixSimulator(n, 0)
    sim.evolve(ins)
    assert np.max(np.abs(sim.density_matrix().ravel(order="F") - exp)) < TOL
    per_op = DensityMatrixSimulator(n, 0)
    per_op.evolve(ins, mode=1)  # the separate block kernel
    assert np.max(np.abs(per_op.density_matrix().ravel(order="F") - exp)) < TOL


def test_dm_apply_gate_and_noise_api():
    n = 5
    rng = np.random.default_rng(77)
    ops = []
    for _ in range(12):
        ops += random_circuit(n, 2, rng)
        ops.append(random_noise(n, rng))
    exp = O.evolve_dm(n, ops)
    sim = DensityMatrixSimulator(n, 0)
    for i in ops:
        if isinstance(i.operator, G.Gate):
            sim.apply_gate(i.operator, *i.target)
        else:
            sim.apply_noise(i.operator, *i.target)
    assert np.max(np.abs(sim.density_matrix().ravel(order="F") - exp)) < TOL


def test_config1_ry_qft_20q():
    """20-qubit RY + textbook QFT: amplitudes vs the oracle and vs the
    closed form sqrt(N) ifft(a); probabilities and every <Z_i>."""
    ins, th = W.ry_qft(20, seed=0)
    exp = O.evolve_sv(20, ins)
    sim = StateVectorSimulator(20, 10000)
    sim.evolve(ins)
    got = sim.state_vector()
    assert np.max(np.abs(got - exp)) < TOL
    assert np.max(np.abs(got - O.qft_state(th))) < TOL
    p = sim.probabilities()
    pe = O.sv_probabilities(exp, 20)
    assert np.max(np.abs(p - pe)) < TOL
    z = sim.expect_z_all()
    ze = [O.sv_expect_pauli(exp, 20, "I" * q + "Z" + "I" * (19 - q)) for q in range(20)]
    assert np.max(np.abs(z - ze)) < TOL


@pytest.mark.parametrize("n", [1, 5, 12, 20])
def test_alias_parallel_bit_exact(n):
    rng = np.random.default_rng(n)
    sim = StateVectorSimulator(n, 0)
    sim.evolve(random_circuit(n, 60, rng))
    p = sim.probabilities()
    wsum = float(p.sum())
    L.call("bsim_alias_build", sim.handle, L.ALIAS_PARALLEL, wsum)
    ap, al = sim.alias_table()
    eap, eal = O.alias_exact(p, wsum)
    np.testing.assert_array_equal(ap, eap)
    np.testing.assert_array_equal(al, eal)
    shots = sim.samples(seed=1234, wsum=wsum, shots=10000)
    idx, u = O.shot_draws(1234, n, 10000)
    np.testing.assert_array_equal(shots, O.lookup(eap, eal.astype(np.int64), idx, u))


@pytest.mark.parametrize("n", [1, 4, 10, 16])
def test_alias_reference_bit_exact(n):
    rng = np.random.default_rng(n + 11)
    sim = StateVectorSimulator(n, 0)
    sim.evolve(random_circuit(n, 60, rng))
    p = sim.probabilities()
    wsum = float(p.sum())
    L.call("bsim_alias_build", sim.handle, L.ALIAS_REFERENCE, wsum)
    ap, al = sim.alias_table()
    eap, eal = O.alias_reference(p, wsum)
    np.testing.assert_array_equal(ap, eap)
    # entries the reference never assigns stay at their initial value; the
    # lookup never reads them (their ap is >= 1)
    assigned = eap < 1.0
    np.testing.assert_array_equal(al[assigned], eal[assigned])
    idx = np.arange(2 ** n, dtype=np.int64).repeat(3)
    u = np.tile([0.0, 0.5, 0.999999], 2 ** n)
    np.testing.assert_array_equal(sim.sample_from_draws(idx, u), O.lookup(eap, eal, idx, u))


def test_bell_samples(fixtures):
    # test/test_sv_simulator.jl:385-398
    sim = StateVectorSimulator(2, 10000)
    sim.evolve([Instruction(G.H(), [0]), Instruction(G.CNot(), [0, 1])])
    s = sim.samples(seed=5)
    keys, counts = np.unique(s, return_counts=True)
    assert list(keys) == [0, 3]
    assert 0.4 < counts[0] / 10000 < 0.6


def test_expectations_vs_oracle():
    n = 8
    rng = np.random.default_rng(21)
    ins = random_circuit(n, 60, rng)
    psi = O.evolve_sv(n, ins)
    sim = StateVectorSimulator(n, 0)
    sim.evolve(ins)
    for word in ["ZIIIIIII", "IXIYIIZI", "XXXXXXXX", "YZYZYZYZ", "IIIIIIIY"]:
        assert abs(sim.expect_pauli_words([word])[0] - O.sv_expect_pauli(psi, n, word)) < TOL
    for k, t in [(1, [3]), (2, [5, 1]), (2, [1, 5]), (3, [6, 0, 2]), (4, [1, 2, 3, 7])]:
        h = random_unitary(k, rng)
        h = h + h.conj().T
        got = sim.expectation(HermitianObservable(h), *t)
        assert abs(got - O.sv_expect_matrix(psi, n, h, t)) < TOL
    got = sim.expectation(HObs(), 4)
    assert abs(got - O.sv_expect_matrix(psi, n, HObs().matrix(), [4])) < TOL
    tp = TensorProduct([XO(), ZO(), YO()])
    assert abs(sim.expectation(tp, 0, 4, 7) - O.sv_expect_pauli(psi, n, "XIIIZIIY")) < TOL
    m = sim.marginal_probabilities([6, 2, 3])
    np.testing.assert_allclose(m, O.marginal(O.sv_probabilities(psi, n), n, [6, 2, 3]), atol=TOL)


def test_dm_expectations_vs_oracle():
    n = 5
    rng = np.random.default_rng(31)
    ops = []
    for _ in range(10):
        ops += random_circuit(n, 2, rng)
        ops.append(random_noise(n, rng))
    rho = O.evolve_dm(n, ops)
    sim = DensityMatrixSimulator(n, 0)
    sim.evolve(ops)
    for word in ["ZZIII", "IXIYI", "XYZXY", "IIIIZ"]:
        assert abs(sim.expect_pauli_words([word])[0] - O.dm_expect_pauli(rho, n, word)) < TOL
    for k, t in [(1, [2]), (2, [4, 1]), (3, [0, 3, 1])]:
        h = random_unitary(k, rng)
        h = h + h.conj().T
        assert abs(sim.expectation(HermitianObservable(h), *t) - O.dm_expect_matrix(rho, n, h, t)) < TOL
    zs = sim.expect_z_all()
    for q in range(n):
        assert abs(zs[q] - O.dm_expect_pauli(rho, n, "I" * q + "Z" + "I" * (n - 1 - q))) < TOL


def test_brickwall_fused_vs_oracle_22q():
    _say("brickwall 22q")
    ins = W.brickwall(22, 12, seed=3)
    exp = O.evolve_sv(22, ins)
    sim = StateVectorSimulator(22, 0)
    sim.evolve(ins)
    assert np.max(np.abs(sim.state_vector() - exp)) < TOL


def test_brickwall_30q_properties():
    """Full size: norm preserved, fused and per-op paths agree on every
    <Z_i>, and 1e6 shots from the parallel alias table stay in range."""
    _say("brickwall 30q")
    ins = W.brickwall(30, 20, seed=0)
    sim = StateVectorSimulator(30, 1000000)
    sim.evolve(ins)
    z_fused = sim.expect_z_all()
    p_total = sim.expect_pauli_words(["I" * 30])[0]
    assert abs(p_total - 1.0) < 1e-10
    shots = sim.samples(seed=7)
    assert shots.min() >= 0 and shots.max() < 2 ** 30
    sim.reinit(30, 0)
    sim.evolve(ins, mode=1)
    z_single = sim.expect_z_all()
    assert np.max(np.abs(z_fused - z_single)) < TOL
    del sim


def test_shots_to_bits():
    sim = StateVectorSimulator(4, 8)
    sim.evolve([Instruction(G.X(), [0]), Instruction(G.X(), [3])])
    res = bsim.simulate(sim, bsim.Program([Instruction(G.X(), [0]), Instruction(G.X(), [3])]), shots=8, seed=3)
    assert res.measurements.shape == (8, 4)
    assert (res.measurements == np.array([1, 0, 0, 1])).all()


def test_errors_are_reported():
    sim = StateVectorSimulator(3, 0)
    with pytest.raises(bsim.BsimError, match="range"):
        sim.apply_gate(G.H(), 7)
    with pytest.raises(bsim.BsimError):
        sim.samples(shots=10, seed=1, alias_mode=99)


# ---- adjoint differentiation ---------------------------------------------------

def _oracle_energy(n, ins, words, coeffs):
    psi = O.evolve_sv(n, ins)
    return sum(c * O.sv_expect_pauli(psi, n, w) for w, c in zip(words, coeffs))


@pytest.mark.parametrize("n,layers", [(4, 2), (8, 3), (12, 4)])
def test_adjoint_vs_parameter_shift(n, layers):
    """Adjoint gradient of the hardware-efficient ansatz vs parameter shift
    through the plain GPU path (1e-10) and vs the oracle's parameter shift."""
    from bsim import adjoint as A

    _say(f"adjoint n={n}")
    ins, pmap, th = W.hea(n, layers, seed=n)
    words, coeffs = W.tfim(n)
    sim = StateVectorSimulator(n, 0)
    ev, grad = A.adjoint_gradient(sim, ins, pmap, words, coeffs)
    assert abs(ev - _oracle_energy(n, ins, words, coeffs)) < TOL
    ps = A.parameter_shift_gradient(sim, ins, pmap, th, words, coeffs)
    assert np.max(np.abs(grad - ps)) < TOL
    if n <= 8:
        ops = []
        for p in range(len(th)):
            tp = th.copy(); tp[p] += math.pi / 2
            tm = th.copy(); tm[p] -= math.pi / 2
            ops.append(0.5 * (_oracle_energy(n, A.with_params(ins, pmap, tp), words, coeffs)
                              - _oracle_energy(n, A.with_params(ins, pmap, tm), words, coeffs)))
        assert np.max(np.abs(grad - np.array(ops))) < TOL
    fd = A.finite_difference_gradient(sim, ins, pmap, th, words, coeffs, h=1e-6)
    assert np.max(np.abs(grad - fd)) < 1e-7  # FD roundoff ~ eps/h


def test_adjoint_general_gates():
    """U(theta, phi, lambda), controlled rotations, ZZ and CPhaseShift
    parameters, checked against central differences of the oracle."""
    from bsim import adjoint as A

    n = 6
    rng = np.random.default_rng(3)
    ins, pmap = [], {}
    p = 0
    for _ in range(5):
        for q in range(n):
            ins.append(Instruction(G.U(*rng.uniform(-3, 3, 3)), [q]))
            for j in range(3):
                pmap[(len(ins) - 1, j)] = p
                p += 1
        for q in range(n - 1):
            g = [G.ZZ(float(rng.uniform(-3, 3))), G.CPhaseShift(float(rng.uniform(-3, 3))),
                 G.Control(G.Rx(float(rng.uniform(-3, 3))), (1,)), G.XX(float(rng.uniform(-3, 3)))][q % 4]
            ins.append(Instruction(g, [q, q + 1]))
            pmap[(len(ins) - 1, 0)] = p
            p += 1
        ins.append(Instruction(G.CNot(), [0, n - 1]))
    th = np.zeros(p)
    for (i, j), k in pmap.items():
        th[k] = ins[i].operator.angles[j]
    words = ["ZZIIII", "IXIXII", "YIIIIZ", "IIZYXI"]
    coeffs = [0.7, -0.3, 1.1, 0.25]
    sim = StateVectorSimulator(n, 0)
    ev, grad = A.adjoint_gradient(sim, ins, pmap, words, coeffs)
    h = 1e-5
    fd = np.zeros(p)
    for k in range(p):
        tp = th.copy(); tp[k] += h
        tm = th.copy(); tm[k] -= h
        fd[k] = (_oracle_energy(n, A.with_params(ins, pmap, tp), words, coeffs)
                 - _oracle_energy(n, A.with_params(ins, pmap, tm), words, coeffs)) / (2 * h)
    assert abs(ev - _oracle_energy(n, ins, words, coeffs)) < TOL
    assert np.max(np.abs(grad - fd)) < 1e-8


def test_adjoint_torch_function():
    import torch
    from bsim import adjoint as A

    n = 5
    ins, pmap, th = W.hea(n, 2, seed=1)
    words, coeffs = W.tfim(n)
    sim = StateVectorSimulator(n, 0)
    theta = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    e = A.ExpectationFunction.apply(theta, sim, ins, pmap, words, coeffs)
    e.backward()
    _, grad = A.adjoint_gradient(sim, ins, pmap, words, coeffs)
    assert np.max(np.abs(theta.grad.numpy() - grad)) < 1e-12
