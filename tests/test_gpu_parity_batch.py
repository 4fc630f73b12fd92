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
