"""bsim_dm_apply_superop: one pass of a 4^k x 4^k Liouville matrix on the
density matrix (the fused form of a DM gate's two passes and of a physical
Kraus channel).

The index convention is pinned on the CPU first: ``np_superop`` (a tensor
contraction written from the definition rho'[r, c] = sum S[r + D c,
r' + D c'] rho[r', c']) must give U rho U^dagger for S = kron(conj U, U),
with U applied by the oracle's state-vector routine.  On the GPU:
  - a random (non-physical) S on a random stored matrix vs np_superop;
  - S from a unitary / from Kraus matrices vs the library's own DM gate and
    physical Kraus paths, and the oracle's physical Kraus channel (1e-12);
  - superoperators inside evolve (fused plan) and in the DM adjoint.
"""
import sys

import numpy as np
import pytest

import oracle as O
from bsim import adjoint as A
from bsim import gates as G
from bsim import noises as Nz
from bsim import workloads as W
from bsim.simulators import DensityMatrixSimulator, Instruction, Superoperator

from circuits import random_unitary

TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def np_superop(rho, n, S, targets):
    """rho'[r, c] = sum S[r + D c, r' + D c'] rho[r', c'] on the physical
    matrix rho (qubit 0 = most significant bit); r, c index the targets'
    block with targets[0] as the most significant bit."""
    k = len(targets)
    T = np.asarray(rho, dtype=np.complex128).reshape([2] * (2 * n))
    Sr = np.asarray(S, dtype=np.complex128).reshape([2] * (4 * k))  # axes: c bits, r bits, c' bits, r' bits
    axes_T = [n + t for t in targets] + list(targets)
    out = np.tensordot(Sr, T, axes=(list(range(2 * k, 4 * k)), axes_T))
    rest = [a for a in range(2 * n) if a not in axes_T]
    out = np.moveaxis(out, list(range(2 * n)), axes_T + rest)
    return out.reshape(2 ** n, 2 ** n)


def _rand_rho(n, rng):
    a = rng.normal(size=(2 ** n, 2 ** n)) + 1j * rng.normal(size=(2 ** n, 2 ** n))
    r = a @ a.conj().T
    return r / np.trace(r)


def _u_rho_udag(rho, n, U, targets):
    N = 2 ** n
    M = np.stack([O.np_apply(rho[:, j].copy(), n, U, targets) for j in range(N)], axis=1)  # U rho
    Md = M.conj().T
    return np.stack([O.np_apply(Md[:, j].copy(), n, U, targets) for j in range(N)], axis=1).conj().T


@pytest.mark.parametrize("targets", [[1], [2, 0], [0, 3, 1]])
def test_np_superop_is_u_rho_udagger(targets):
    n = 4
    rng = np.random.default_rng(len(targets))
    U = random_unitary(len(targets), rng)
    rho = _rand_rho(n, rng)
    got = np_superop(rho, n, np.kron(U.conj(), U), targets)
    np.testing.assert_allclose(got, _u_rho_udag(rho, n, U, targets), atol=1e-13)


def _upload_physical(sim, rho):
    sim.upload(rho.T)  # storage is rho^T


def _physical(sim):
    return sim.density_matrix().T


@pytest.mark.gpu
@pytest.mark.parametrize("n,targets", [(3, [1]), (5, [4, 0]), (5, [1, 3]), (6, [5, 2, 0]), (6, [0, 1, 2]), (7, [3, 6, 1])])
def test_superop_random_matrix(n, targets):
    rng = np.random.default_rng(n * 7 + len(targets))
    k = len(targets)
    S = rng.normal(size=(4 ** k, 4 ** k)) + 1j * rng.normal(size=(4 ** k, 4 ** k))
    rho = _rand_rho(n, rng)
    _say(f"superop random n={n} targets={targets}")
    sim = DensityMatrixSimulator(n, 0)
    _upload_physical(sim, rho)
    sim.apply_superop(S, *targets)
    np.testing.assert_allclose(_physical(sim), np_superop(rho, n, S, targets), atol=1e-11)


@pytest.mark.gpu
@pytest.mark.parametrize("targets", [[2], [0, 3], [4, 1, 2]])
def test_superop_unitary_matches_dm_gate_and_kraus(targets):
    n = 5
    rng = np.random.default_rng(sum(targets) + 11)
    k = len(targets)
    U = random_unitary(k, rng)
    pre = [Instruction(G.H(), [q]) for q in range(n)] + [Instruction(G.Rx(0.4 * q + 0.1), [q]) for q in range(n)]
    pre += [Instruction(G.CNot(), [q, q + 1]) for q in range(n - 1)]
    _say(f"superop unitary targets={targets}")
    a = DensityMatrixSimulator(n, 0).evolve(pre)
    a.apply_superop(Superoperator.from_unitary(U), *targets)
    Ub = U
    if k > 2:  # the reference's k > 2 Kraus matrices index the sorted targets
        perm = [list(targets).index(t) for t in sorted(targets)]
        Ub = U.reshape([2] * (2 * k)).transpose(perm + [k + p for p in perm]).reshape(2 ** k, 2 ** k)
    b = DensityMatrixSimulator(n, 0, noise_mode="physical").evolve(pre)
    b.apply_noise(Nz.Kraus([Ub]), *targets)
    np.testing.assert_allclose(a.density_matrix(), b.density_matrix(), atol=TOL)
    rho0 = O.evolve_dm(n, pre, "physical").reshape(2 ** n, 2 ** n, order="F").T
    np.testing.assert_allclose(_physical(a), _u_rho_udag(rho0, n, U, targets), atol=TOL)
    if k == 1:
        c = DensityMatrixSimulator(n, 0).evolve(pre)
        c.apply_gate(G.Unitary(U), *targets)
        np.testing.assert_allclose(a.density_matrix(), c.density_matrix(), atol=TOL)


@pytest.mark.gpu
@pytest.mark.parametrize("noise,targets", [(Nz.Depolarizing(0.1), [3]), (Nz.AmplitudeDamping(0.3), [0]),
                                           (Nz.TwoQubitDepolarizing(0.05), [1, 4]),
                                           (Nz.TwoQubitDephasing(0.2), [3, 2])])
def test_superop_from_kraus_vs_oracle(noise, targets):
    n = 5
    pre = [Instruction(G.H(), [q]) for q in range(n)] + [Instruction(G.Ry(0.3 * q - 0.5), [q]) for q in range(n)]
    pre += [Instruction(G.CZ(), [q, q + 1]) for q in range(n - 1)]
    _say(f"superop kraus {noise!r} {targets}")
    sim = DensityMatrixSimulator(n, 0).evolve(pre)
    sim.apply_superop(Superoperator.from_kraus(noise.kraus_rep()), *targets)
    ref = O.evolve_dm(n, pre + [Instruction(noise, targets)], "physical")
    np.testing.assert_allclose(sim.density_matrix().ravel(order="F"), ref, atol=TOL)


@pytest.mark.gpu
def test_superop_in_evolve_and_adjoint():
    n = 4
    ins, pmap, th = W.hea(n, 2, seed=4)
    rng = np.random.default_rng(3)
    ks = Nz.TwoQubitDepolarizing(0.07).kraus_rep()
    sop = Superoperator.from_kraus(ks)
    ins_s = ins + [Instruction(sop, [0, 2]), Instruction(Superoperator.from_unitary(random_unitary(1, rng)), [3])]
    _say("superop in evolve / adjoint")
    a = DensityMatrixSimulator(n, 0).evolve(ins_s)
    b = DensityMatrixSimulator(n, 0).evolve(ins)
    b.apply_superop(sop, 0, 2)
    b.apply_superop(ins_s[-1].operator, 3)
    np.testing.assert_allclose(a.density_matrix(), b.density_matrix(), atol=TOL)
    # adjoint through a superoperator: vs the same circuit with the physical Kraus channel
    words, coeffs = W.tfim(n)
    ins_s2 = ins + [Instruction(sop, [0, 2])] + ins
    ins_k2 = ins + [Instruction(Nz.TwoQubitDepolarizing(0.07), [0, 2])] + ins
    pmap2 = dict(pmap)
    off = len(ins) + 1
    for (i, j), p in pmap.items():
        pmap2[(i + off, j)] = p
    ev_s, g_s = A.adjoint_gradient(DensityMatrixSimulator(n, 0), ins_s2, pmap2, words, coeffs)
    ev_k, g_k = A.adjoint_gradient(DensityMatrixSimulator(n, 0, noise_mode="physical"), ins_k2, pmap2, words, coeffs)
    assert abs(ev_s - ev_k) < 1e-11
    np.testing.assert_allclose(g_s, g_k, atol=1e-11)
