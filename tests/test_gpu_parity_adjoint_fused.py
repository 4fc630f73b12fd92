"""The fused backward sweep of the state-vector adjoint (tile passes over the
reversed U^dagger ops, psi and lambda both in LDS, parameter products
reduced inside the pass) against the one-kernel-per-op backward sweep and
the oracle's parameter shift.

The backward plan itself is pinned on the CPU (tests/test_planner.py
interprets it against the oracle); here the tile kernel runs it: the two
device paths must agree to summation order (1e-10), the fused gradient must
be bit-identical from run to run, and sizes cover one tile (n <= 11), many
tiles, and the sub-cube threads a small tile leaves idle (n = 5..8)."""
import sys

import numpy as np
import pytest

import oracle as O
from bsim import adjoint as A
from bsim import gates as G
from bsim import workloads as W
from bsim.simulators import Instruction, StateVectorSimulator

from circuits import random_unitary

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _both(n, ins, pmap, words, coeffs):
    sim = StateVectorSimulator(n, 0)
    ev_f, g_f = A.adjoint_gradient(sim, ins, pmap, words, coeffs, fused=True)
    psi_f = sim.state_vector().copy()
    ev_p, g_p = A.adjoint_gradient(sim, ins, pmap, words, coeffs, fused=False)
    np.testing.assert_array_equal(psi_f, sim.state_vector())  # the forward pass is the same
    return ev_f, g_f, ev_p, g_p


@pytest.mark.parametrize("n,layers", [(5, 2), (8, 3), (11, 2), (14, 2), (17, 1)])
def test_fused_adjoint_matches_per_op(n, layers):
    _say(f"fused adjoint hea n={n} layers={layers}")
    ins, pmap, th = W.hea(n, layers, seed=100 + n)
    words, coeffs = W.tfim(n)
    ev_f, g_f, ev_p, g_p = _both(n, ins, pmap, words, coeffs)
    assert ev_f == ev_p
    np.testing.assert_allclose(g_f, g_p, atol=TOL)


def test_fused_adjoint_vs_oracle_parameter_shift():
    n = 9
    _say("fused adjoint vs oracle parameter shift n=9")
    ins, pmap, th = W.hea(n, 2, seed=7)
    words, coeffs = W.tfim(n)
    _, grad = A.adjoint_gradient(StateVectorSimulator(n, 0), ins, pmap, words, coeffs, fused=True)

    def energy(t):
        psi = O.evolve_sv(n, A.with_params(ins, pmap, t))
        return sum(c * O.sv_expect_pauli(psi, n, w) for w, c in zip(words, coeffs))

    ps = np.zeros(len(th))
    for p in range(len(th)):
        tp = th.copy(); tp[p] += np.pi / 2
        tm = th.copy(); tm[p] -= np.pi / 2
        ps[p] = 0.5 * (energy(tp) - energy(tm))
    np.testing.assert_allclose(grad, ps, atol=TOL)


@pytest.mark.parametrize("n", [7, 13])
def test_fused_adjoint_general_gates(n):
    """Three-parameter U, diagonal ZZ / CPhaseShift, controlled rotations
    with either control value and a control outside the tile (n = 13), XX,
    Rx(0) and a parameter-free 3-qubit unitary (a non-tile step)."""
    _say(f"fused adjoint general gates n={n}")
    rng = np.random.default_rng(90 + n)
    ins, pmap = [], {}
    p = 0
    for layer in range(3):
        for q in range(n):
            ins.append(Instruction(G.U(*rng.uniform(-3, 3, 3)), [q]))
            for j in range(3):
                pmap[(len(ins) - 1, j)] = p
                p += 1
        ins.append(Instruction(G.Rx(0.0), [layer]))
        pmap[(len(ins) - 1, 0)] = p
        p += 1
        for q in range(n - 1):
            a = float(rng.uniform(-3, 3))
            g = [G.ZZ(a), G.CPhaseShift(a), G.Control(G.Rx(a), (1,)), G.XX(a), G.Control(G.Ry(a), (0,))][q % 5]
            t = [q, q + 1] if q % 2 == 0 else [(q + 5) % n, q + 1]
            if t[0] == t[1]:
                t = [q, q + 1]
            ins.append(Instruction(g, t))
            pmap[(len(ins) - 1, 0)] = p
            p += 1
        ins.append(Instruction(G.Unitary(random_unitary(3, rng)), [0, n // 2, n - 1]))
    words, coeffs = W.tfim(n)
    words = list(words) + ["Y" + "I" * (n - 2) + "X"]
    coeffs = list(coeffs) + [0.3]
    ev_f, g_f, ev_p, g_p = _both(n, ins, pmap, words, coeffs)
    assert ev_f == ev_p
    np.testing.assert_allclose(g_f, g_p, atol=TOL)


def test_fused_adjoint_deterministic():
    n = 15
    _say("fused adjoint determinism n=15")
    ins, pmap, th = W.hea(n, 2, seed=3)
    words, coeffs = W.tfim(n)
    sim = StateVectorSimulator(n, 0)
    _, g1 = A.adjoint_gradient(sim, ins, pmap, words, coeffs)
    _, g2 = A.adjoint_gradient(sim, ins, pmap, words, coeffs)
    np.testing.assert_array_equal(g1, g2)


def test_fused_adjoint_shared_parameters_and_small_state():
    """One parameter on several gates (the products add), and a 4-qubit
    state (below the fused path's minimum: the per-op sweep runs)."""
    for n in (4, 10):
        _say(f"fused adjoint shared parameters n={n}")
        ins, pmap, th = W.hea(n, 2, seed=n)
        pm = {k: v % 3 for k, v in pmap.items()}
        words, coeffs = W.tfim(n)
        ev_f, g_f, ev_p, g_p = _both(n, ins, pm, words, coeffs)
        assert ev_f == ev_p
        np.testing.assert_allclose(g_f, g_p, atol=TOL)


def test_fused_adjoint_without_parameters():
    """No parametrized gate: the backward sweep carries no products, the
    gradient is empty and <H> matches the oracle."""
    n = 12
    _say("fused adjoint without parameters n=12")
    ins, _, _ = W.hea(n, 2, seed=1)
    words, coeffs = W.tfim(n)
    ev, grad = A.adjoint_gradient(StateVectorSimulator(n, 0), ins, {}, words, coeffs, n_params=0, fused=True)
    psi = O.evolve_sv(n, ins)
    assert grad.shape == (0,)
    assert abs(ev - sum(c * O.sv_expect_pauli(psi, n, w) for w, c in zip(words, coeffs))) < TOL


@pytest.mark.parametrize("n,layers,seed", [(7, 2, 0), (12, 3, 1), (14, 2, 2)])
def test_fused_adjoint_merged_runs(n, layers, seed):
    """Merged one-qubit runs of every shape (Pauli-coded products before and
    after the merged op, payload products after it, runs ended by a
    parameter-free op, a non-Pauli payload or a controlled op): fused vs
    per-op, and at n = 7 vs the oracle's central differences."""
    from circuits import one_qubit_runs

    _say(f"fused adjoint merged runs n={n} layers={layers} seed={seed}")
    ins, pmap, th = one_qubit_runs(n, layers, seed)
    words, coeffs = W.tfim(n)
    words = list(words) + ["Y" + "I" * (n - 2) + "X"]
    coeffs = list(coeffs) + [0.3]
    ev_f, g_f, ev_p, g_p = _both(n, ins, pmap, words, coeffs)
    assert ev_f == ev_p
    np.testing.assert_allclose(g_f, g_p, atol=TOL)
    if n == 7:
        def energy(t):
            psi = O.evolve_sv(n, A.with_params(ins, pmap, t))
            return sum(c * O.sv_expect_pauli(psi, n, w) for w, c in zip(words, coeffs))

        hstep = 1e-5
        fd = np.zeros(len(th))
        for k in range(len(th)):
            tp = th.copy(); tp[k] += hstep
            tm = th.copy(); tm[k] -= hstep
            fd[k] = (energy(tp) - energy(tm)) / (2 * hstep)
        np.testing.assert_allclose(g_f, fd, atol=1e-8)


def _mixed_words(n, rng, gathered=True):
    """Pauli words of every shape lambda = H psi plans for: Z Z and X terms
    (diagonal group, single flips in the first tile and in later ones), Y X
    across the low and high bits, two-flip words within the low bits, and
    with `gathered` a word flipping every qubit (no tile holds it: the
    gather kernel, after which <psi|lambda> is its own reduction) and random
    words with random coefficients."""
    words, coeffs = W.tfim(n)
    words, coeffs = list(words), list(coeffs)
    words.append("Y" + "I" * (n - 2) + "X")
    coeffs.append(0.3)
    words.append("I" * (n - 4) + "XZYI")
    coeffs.append(-0.35)
    if gathered:
        words.append("X" * n)
        coeffs.append(-0.2)
        words.append("Y" * (n // 2) + "Z" * (n - n // 2))
        coeffs.append(0.15)
        for _ in range(6):
            words.append("".join(rng.choice(list("IXYZ"), n)))
            coeffs.append(float(rng.uniform(-1, 1)))
    return words, coeffs


@pytest.mark.parametrize("n,gathered", [(11, True), (12, True), (13, True), (16, True), (13, False), (16, False)])
def test_lambda_tile_passes_vs_oracle(n, gathered):
    """lambda = H psi as tile passes (words grouped by flip mask, each group
    added in the first pass whose LDS tile of psi holds its flips; the rest
    by gathers; <H> from the last tile pass when no gather follows): the
    adjoint's <H> = Re <psi|lambda> against the oracle's sum of word
    expectations, and the gradient (which reads all of lambda) against the
    oracle's parameter shift, exact for RY and RZ, on every parameter
    (n <= 13) or a sample of 12 (n = 16)."""
    _say(f"lambda tile passes n={n} gathered={gathered}")
    rng = np.random.default_rng(500 + n)
    ins, pmap, th = W.hea(n, 2, seed=60 + n)
    words, coeffs = _mixed_words(n, rng, gathered)
    ev, grad = A.adjoint_gradient(StateVectorSimulator(n, 0), ins, pmap, words, coeffs, fused=True)

    def energy(t):
        psi = O.evolve_sv(n, A.with_params(ins, pmap, t))
        return sum(c * O.sv_expect_pauli(psi, n, w) for w, c in zip(words, coeffs))

    assert abs(ev - energy(th)) < TOL
    sample = range(len(th)) if n <= 13 else rng.choice(len(th), 12, replace=False)
    for p in sample:
        tp = th.copy(); tp[p] += np.pi / 2
        tm = th.copy(); tm[p] -= np.pi / 2
        assert abs(grad[p] - 0.5 * (energy(tp) - energy(tm))) < TOL, p
