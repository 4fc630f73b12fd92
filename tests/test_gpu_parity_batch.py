"""GPU parity of the persistent batch launch (bsim_batch_run) against the
oracle and against the per-circuit path.

Bar: <Z_q> within 1e-10 of the oracle's state; shot indices bit-exact
against the oracle's inverse-CDF rule applied to the same final amplitudes.
The batch plans and tile passes are those of bsim_evolve, so the state the
per-circuit path downloads is the batch's state bit for bit.
"""
import sys

import numpy as np
import pytest

import oracle as O
import bsim
from bsim import _lib as L
from bsim import gates as G
from bsim import noises as Nz
from bsim import workloads as W
from bsim.batch import run_batch
from bsim.simulators import Instruction, StateVectorSimulator

from circuits import random_circuit

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _z_of(psi, n):
    p = np.abs(psi) ** 2
    idx = np.arange(2 ** n)
    return np.array([p[((idx >> (n - 1 - q)) & 1) == 0].sum() - p[((idx >> (n - 1 - q)) & 1) == 1].sum()
                     for q in range(n)])


def _gpu_state(n, ins):
    sim = StateVectorSimulator(n, 0)
    sim.evolve(ins)
    return sim.state_vector()


def test_batch_small_vs_oracle():
    rng = np.random.default_rng(11)
    circuits = []
    for i, n in enumerate([4, 5, 7, 8, 9, 10, 12, 13, 15, 16]):
        if i % 2:
            ins = W.brickwall(n, int(rng.integers(2, 12)), seed=100 + i)
        else:
            ins = random_circuit(n, 40, rng, kinds=("1q", "2q", "ctrl", "diag", "pow"))
        circuits.append((n, ins))
    _say(f"batch of {len(circuits)} circuits, n = {[c[0] for c in circuits]}")
    res = run_batch(circuits, shots=1000, seeds=[7 + i for i in range(len(circuits))])
    for i, ((n, ins), r) in enumerate(zip(circuits, res)):
        exp = O.evolve_sv(n, ins)
        assert np.max(np.abs(r.expect_z - _z_of(exp, n))) < TOL, (i, n)
        psi = _gpu_state(n, ins)
        assert np.max(np.abs(psi - exp)) < TOL
        np.testing.assert_array_equal(r.samples, O.cdf_shots(psi, n, 1000, 7 + i), err_msg=f"circuit {i} n={n}")
        assert np.isfinite(r.latency_us) and r.latency_us > 0


def test_batch_uneven_many():
    """200 brick-wall circuits, n in 8..18, depth 10..40, against the
    per-circuit launches of the same plans."""
    spec = W.batch_tasks(200, seed=5, qmin=8, qmax=18)
    circuits = [(n, W.brickwall(n, d, seed=s)) for n, d, s in spec]
    _say(f"uneven batch of {len(circuits)}")
    res = run_batch(circuits, shots=1000, seeds=[s for _, _, s in spec])
    for (n, ins), (_, _, s), r in zip(circuits, spec, res):
        psi = _gpu_state(n, ins)
        assert np.max(np.abs(r.expect_z - _z_of(psi, n))) < 1e-12
        np.testing.assert_array_equal(r.samples, O.cdf_shots(psi, n, 1000, s))
    lat = np.array([r.latency_us for r in res])
    assert (lat > 0).all() and np.isfinite(lat).all()


def test_batch_edge_cases():
    """No gates, zero shots, a 22-qubit circuit beside 4-qubit ones, and one
    circuit repeated with different seeds."""
    one = [Instruction(G.H(), [0]), Instruction(G.CNot(), [0, 1])]
    circuits = [(4, []), (4, one), (22, W.brickwall(22, 3, seed=9)), (6, one), (6, one)]
    _say("batch edge cases")
    res = run_batch(circuits, shots=[10, 0, 500, 2000, 2000], seeds=[1, 2, 3, 4, 5])
    np.testing.assert_allclose(res[0].expect_z, np.ones(4), atol=TOL)
    assert (res[0].samples == 0).all() and len(res[0].samples) == 10
    np.testing.assert_allclose(res[1].expect_z, [0, 0, 1, 1], atol=TOL)
    assert len(res[1].samples) == 0
    psi = _gpu_state(22, circuits[2][1])
    assert np.max(np.abs(res[2].expect_z - _z_of(psi, 22))) < 1e-12
    np.testing.assert_array_equal(res[2].samples, O.cdf_shots(psi, 22, 500, 3))
    for r, seed in ((res[3], 4), (res[4], 5)):
        assert set(np.unique(r.samples)) <= {0, 48}
        np.testing.assert_array_equal(r.samples, O.cdf_shots(_gpu_state(6, one), 6, 2000, seed))
    assert not np.array_equal(res[3].samples, res[4].samples)


def test_batch_rejects():
    with pytest.raises(bsim.BsimError, match="4..26"):
        run_batch([(3, [Instruction(G.H(), [0])])], shots=1)
    u3 = np.linalg.qr(np.random.default_rng(0).normal(size=(8, 8)) + 0j)[0]
    with pytest.raises(bsim.BsimError, match="fusable"):
        run_batch([(5, [Instruction(G.Unitary(u3), [0, 1, 2])])], shots=1)
    with pytest.raises(TypeError):
        run_batch([(5, [Instruction(Nz.BitFlip(0.1), [0])])], shots=1)


def test_simulate_batch_persistent():
    """simulate_batch over the persistent launch: measurement rows are the
    shot indices' bits of the measured qubits."""
    progs = [bsim.Program([Instruction(G.H(), [0]), Instruction(G.CNot(), [0, 1]), Instruction(G.X(), [4])])
             for _ in range(3)]
    sim = StateVectorSimulator(5, 100)
    out = bsim.simulate_batch(sim, progs, shots=100, seed=3)
    for r in out:
        assert r.measurements.shape == (100, 5)
        rows = {tuple(x) for x in r.measurements}
        assert rows <= {(0, 0, 0, 0, 1), (1, 1, 0, 0, 1)} and len(rows) == 2


def test_batch_gradient_sweep_matches_serial():
    """Parameter sets of the hardware-efficient ansatz (5..16 qubits, random
    angles) as gradient tasks in the same launch as ordinary circuits: every
    batched gradient equals bsim_adjoint_gradient on the same circuit within
    1e-10 (the backward plans and per-wave sums are the same), <H> too, and
    the circuits' results are unchanged by the company.  Odd tasks pass op
    lists lowered once (with the parameter map) to both paths."""
    from bsim import adjoint as A
    from bsim.batch import GradientTask
    from bsim.simulators import lower_instructions

    items, serial = [], []
    for k, n in enumerate([5, 6, 8, 11, 12, 13, 16, 16]):
        ins, pmap, th = W.hea(n, 3 if n < 16 else 8, seed=40 + k)
        words, coeffs = W.tfim(n, hx=0.3 + 0.1 * k)
        src = lower_instructions(list(ins), L.BSIM_SV, param_map=pmap) if k % 2 else ins
        items.append(GradientTask(n, src, pmap, words, coeffs))
        sim = StateVectorSimulator(n, 0)
        serial.append(A.adjoint_gradient(sim, src, pmap, words, coeffs))
    spec = W.batch_tasks(40, seed=3, qmin=8, qmax=16)
    circuits = [(n, W.brickwall(n, d, seed=s)) for n, d, s in spec]
    mixed = []
    for i in range(max(len(items), len(circuits))):
        if i < len(circuits):
            mixed.append(circuits[i])
        if i < len(items):
            mixed.append(items[i])
    seeds = list(range(len(mixed)))
    _say(f"batch: {len(items)} gradient tasks among {len(circuits)} circuits")
    res = run_batch(mixed, shots=[0 if isinstance(m, GradientTask) else 500 for m in mixed], seeds=seeds)
    gi = ci = 0
    for m, r, sd in zip(mixed, res, seeds):
        if isinstance(m, GradientTask):
            ev, g = serial[gi]
            assert abs(r.expval - ev) < TOL, gi
            np.testing.assert_allclose(r.grad, g, atol=TOL, err_msg=f"gradient task {gi}")
            assert r.latency_us > 0
            gi += 1
        else:
            n, ins = m
            psi = _gpu_state(n, ins)
            assert np.max(np.abs(r.expect_z - _z_of(psi, n))) < 1e-12
            np.testing.assert_array_equal(r.samples, O.cdf_shots(psi, n, 500, sd))
            ci += 1
    assert gi == len(items) and ci == len(circuits)


def test_batch_gradient_vs_oracle_parameter_shift():
    """A 7-qubit parameter set with controlled rotations and diagonal
    gates: the batched gradient against the oracle's parameter shift."""
    from bsim import adjoint as A
    from bsim.batch import GradientTask

    n = 7
    rng = np.random.default_rng(8)
    ins, pmap, th = [], {}, rng.uniform(-3, 3, 12)
    for p in range(12):
        q = p % n
        g = [G.Rx, G.Ry, G.Rz][p % 3](float(th[p]))
        pmap[(len(ins), 0)] = p
        ins.append(Instruction(g, [q]))
        ins.append(Instruction(G.CNot(), [q, (q + 3) % n]))
        if p % 4 == 0:
            ins.append(Instruction(G.CZ(), [(q + 1) % n, (q + 5) % n]))
    words, coeffs = W.tfim(n)
    _say("batch gradient vs oracle parameter shift")
    r = run_batch([GradientTask(n, ins, pmap, words, coeffs)], shots=0)[0]

    def energy(t):
        psi = O.evolve_sv(n, A.with_params(ins, pmap, t))
        return sum(c * O.sv_expect_pauli(psi, n, w) for w, c in zip(words, coeffs))

    assert abs(r.expval - energy(th)) < TOL
    for p in range(12):
        tp, tm = th.copy(), th.copy()
        tp[p] += np.pi / 2
        tm[p] -= np.pi / 2
        assert abs(r.grad[p] - 0.5 * (energy(tp) - energy(tm))) < TOL, p


def test_batch_gradient_rejects():
    from bsim.batch import GradientTask

    ins, pmap, _ = W.hea(4, 1, seed=0)
    words, coeffs = W.tfim(4)
    with pytest.raises(bsim.BsimError, match="5..26"):
        run_batch([GradientTask(4, ins, pmap, words, coeffs)], shots=0)
    ins, pmap, _ = W.hea(6, 1, seed=0)
    words, coeffs = W.tfim(6)
    u3 = np.linalg.qr(np.random.default_rng(0).normal(size=(8, 8)) + 0j)[0]
    with pytest.raises(bsim.BsimError, match="fusable"):
        run_batch([GradientTask(6, ins + [Instruction(G.Unitary(u3), [0, 1, 2])], pmap, words, coeffs)], shots=0)
    with pytest.raises(bsim.BsimError, match="param_index"):
        run_batch([GradientTask(6, ins, pmap, words, coeffs, n_params=3)], shots=0)
