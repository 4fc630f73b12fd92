"""GPU parity of the density-matrix adjoint (bsim_adjoint_gradient on a
DensityMatrixSimulator): checkpointed forward, segment recomputation,
backward propagation of Lambda through the adjoint maps.

Checks, 1e-10 absolute on gradients and <H>:
  - noise-free circuits: the same gradients as the state-vector adjoint;
  - noisy circuits (both noise semantics): the oracle's parameter-shift
    gradients (exact for Pauli rotations, since every noise map is linear and
    independent of the angle);
  - every checkpoint spacing gives the same result;
  - controlled rotations: oracle central differences (1e-7).
"""
import math
import sys

import numpy as np
import pytest

import oracle as O
from bsim import adjoint as A
from bsim import gates as G
from bsim import noises as Nz
from bsim import workloads as W
from bsim.simulators import DensityMatrixSimulator, Instruction, StateVectorSimulator

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_energy(n, ins, words, coeffs, mode):
    rho = O.evolve_dm(n, ins, mode)
    return sum(c * O.dm_expect_pauli(rho, n, w) for w, c in zip(words, coeffs))


def _oracle_ps(n, ins, pmap, th, words, coeffs, mode):
    g = np.zeros(len(th))
    for p in range(len(th)):
        tp = th.copy()
        tp[p] += math.pi / 2
        tm = th.copy()
        tm[p] -= math.pi / 2
        g[p] = 0.5 * (_oracle_energy(n, A.with_params(ins, pmap, tp), words, coeffs, mode)
                      - _oracle_energy(n, A.with_params(ins, pmap, tm), words, coeffs, mode))
    return g


def test_dm_adjoint_matches_sv_without_noise():
    n = 5
    ins, pmap, th = W.hea(n, 3, seed=2)
    words, coeffs = W.tfim(n)
    _say(f"dm adjoint noise-free {n}q")
    ev_sv, g_sv = A.adjoint_gradient(StateVectorSimulator(n, 0), ins, pmap, words, coeffs)
    ev_dm, g_dm = A.adjoint_gradient(DensityMatrixSimulator(n, 0), ins, pmap, words, coeffs)
    assert abs(ev_sv - ev_dm) < TOL
    assert np.max(np.abs(g_sv - g_dm)) < TOL


@pytest.mark.parametrize("mode", ["reference", "physical"])
def test_dm_adjoint_noisy_vs_oracle_parameter_shift(mode):
    n = 4
    ins, pmap, th = W.hea(n, 2, seed=5, noise_p=0.05)
    ins = ins + [Instruction(Nz.AmplitudeDamping(0.1), [q]) for q in range(n)] + [Instruction(Nz.PhaseFlip(0.07), [1])]
    words, coeffs = W.tfim(n)
    _say(f"dm adjoint noisy {n}q mode={mode}")
    sim = DensityMatrixSimulator(n, 0, noise_mode=mode) if mode != "reference" else DensityMatrixSimulator(n, 0)
    ev, g = A.adjoint_gradient(sim, ins, pmap, words, coeffs)
    assert abs(ev - _oracle_energy(n, ins, words, coeffs, mode)) < TOL
    np.testing.assert_allclose(g, _oracle_ps(n, ins, pmap, th, words, coeffs, mode), atol=TOL)
    # the handle keeps the final state
    np.testing.assert_allclose(sim.density_matrix().ravel(order="F"), O.evolve_dm(n, ins, mode), atol=TOL)


@pytest.mark.parametrize("every", [1, 2, 7, 0, 10000])
def test_dm_adjoint_checkpoint_spacing(every):
    n = 4
    ins, pmap, th = W.hea(n, 2, seed=9, noise_p=0.02)
    words, coeffs = W.tfim(n)
    sim = DensityMatrixSimulator(n, 0)
    ev0, g0 = A.adjoint_gradient(sim, ins, pmap, words, coeffs, checkpoint_every=0)
    ev, g = A.adjoint_gradient(sim, ins, pmap, words, coeffs, checkpoint_every=every)
    assert abs(ev - ev0) < 1e-13
    assert np.max(np.abs(g - g0)) < 1e-13


def test_dm_adjoint_controlled_rotation_vs_differences():
    n = 3
    th = np.array([0.3, -1.1, 0.7, 2.0])
    ins = [Instruction(G.H(), [0]), Instruction(G.Ry(th[0]), [1]),
           Instruction(G.Control(G.Ry(th[1]), (1,)), [0, 2]), Instruction(Nz.Depolarizing(0.03), [2]),
           Instruction(G.Control(G.Rx(th[2]), (0,)), [1, 0]), Instruction(G.CNot(), [2, 1]),
           Instruction(G.Rz(th[3]), [2]), Instruction(Nz.AmplitudeDamping(0.2), [0])]
    pmap = {(1, 0): 0, (2, 0): 1, (4, 0): 2, (6, 0): 3}
    words = ["ZZI", "IXY", "XIZ"]
    coeffs = [1.0, -0.4, 0.25]
    _say("dm adjoint controlled rotations")
    ev, g = A.adjoint_gradient(DensityMatrixSimulator(n, 0), ins, pmap, words, coeffs)
    h = 1e-5
    fd = np.zeros(4)
    for p in range(4):
        tp, tm = th.copy(), th.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (_oracle_energy(n, A.with_params(ins, pmap, tp), words, coeffs, "reference")
                 - _oracle_energy(n, A.with_params(ins, pmap, tm), words, coeffs, "reference")) / (2 * h)
    assert abs(ev - _oracle_energy(n, ins, words, coeffs, "reference")) < TOL
    np.testing.assert_allclose(g, fd, atol=1e-7)
