"""GPU parity of the register chains (tile op kinds 8..12 of the lean
streaming passes) against the oracle: chains of up to four matrices in one
op, longer chains, per-thread signs carried on a chain, the sign ops a
rewrite leaves at the end of a sweep, chains in density-matrix passes
around Kraus passes, and the same plans captured in a graph and run
without allocation.  The CPU planner tests check which forms each circuit
plans into; here the kernels run them.  Bar: 1e-10 absolute."""
import sys

import numpy as np
import pytest

import oracle as O
from bsim import _lib as L
from bsim import workloads as W
from bsim.simulators import DensityMatrixSimulator, StateVectorSimulator, lower_instructions

from circuits import chain_circuit, dm_sparse_noise

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,layers,seed,hot", [(13, 9, 0, 2), (12, 9, 0, 1), (13, 12, 3, 3), (14, 10, 5, 2)])
def test_chain_circuits_vs_oracle(n, layers, seed, hot):
    ins = chain_circuit(n, layers, seed, hot)
    _say(f"chains sv n={n} layers={layers} seed={seed} hot={hot}")
    exp = O.evolve_sv(n, ins)
    for mode in (0, 1):
        got = StateVectorSimulator(n, 0).evolve(ins, mode=mode).state_vector()
        assert np.max(np.abs(got - exp)) < TOL, f"mode {mode}"


@pytest.mark.parametrize("seed", range(6))
def test_chain_random_vs_oracle(seed):
    n = 12 + seed % 3
    ins = chain_circuit(n, 6 + seed, 10 + seed, 1 + seed % 3)
    _say(f"chains sv random n={n} seed={seed}")
    got = StateVectorSimulator(n, 0).evolve(ins).state_vector()
    assert np.max(np.abs(got - O.evolve_sv(n, ins))) < TOL


@pytest.mark.parametrize("mode", ["reference", "physical"])
def test_chain_density_matrix_vs_oracle(mode):
    for n, ins in ((6, W.brickwall(6, 8, seed=6)), (7, dm_sparse_noise(7, 6, 7))):
        _say(f"chains dm n={n} {mode}")
        sim = DensityMatrixSimulator(n, 0, noise_mode=mode).evolve(ins)
        exp = O.evolve_dm(n, ins, mode)
        got = sim.density_matrix()
        assert np.max(np.abs(got.ravel(order="F") - exp)) < TOL


def test_chain_plans_captured_and_without_allocation():
    n = 13
    ins = chain_circuit(n, 9, 0, 2)
    _say("chains captured graph and noalloc n=13")
    eager = StateVectorSimulator(n, 0).evolve(ins).state_vector()
    assert np.max(np.abs(eager - O.evolve_sv(n, ins))) < TOL
    sim = StateVectorSimulator(n, 0)
    with sim.capture(ins, reset=True) as g:
        for rep in range(3):
            g.launch()
            assert np.array_equal(sim.state_vector(), eager), f"replay {rep}"
    b = StateVectorSimulator(n, 0)
    ol = lower_instructions(ins, L.BSIM_SV)
    L.call("bsim_evolve_noalloc", b.handle, ol.ptr, ol.n)
    assert np.array_equal(b.state_vector(), eager)
