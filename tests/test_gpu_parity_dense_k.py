"""Dense k-qubit gates (k = 3..6): 4 to 6 targets run on the FP64 matrix
cores (k_dense_mfma, k_dense_mfma6: the gate's real form as a GEMM over
amplitude groups), 3 through LDS.  Against the oracle at 1e-12 on random unitaries with
random (unsorted) targets, with and without controls, in the per-gate path,
as non-tile steps of a fused evolve, and on density matrices; sizes from one
partial block of groups to many."""
import sys

import numpy as np
import pytest

import oracle as O
from bsim import gates as G
from bsim.simulators import DensityMatrixSimulator, Instruction, StateVectorSimulator

from circuits import random_unitary

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _circuit(n, k, rng, controls):
    """The circuit and its expected state.  A k >= 3 target gate acts with
    its matrix bits on the targets in sorted order (src/gate_kernels.jl:759-764).
    A controlled one (Control of a k-qubit Unitary) keeps matrix bit b on its
    b-th target, and the controls select the subspace; it is checked with numpy
    against the full unitary on controls + targets (MSB first), since the
    reference has no kernel for that arity."""
    ins = []
    psi = None
    def step(instr):
        nonlocal psi
        ins.append(instr)
        psi = O.evolve_sv(n, [instr], psi)
    for q in range(n):
        step(Instruction(G.U(*rng.uniform(0, 6, 3)), [q]))
    for _ in range(3):
        q = [int(x) for x in rng.choice(n, k + controls, replace=False)]
        m = random_unitary(k, rng)
        if controls:
            bits = tuple(int(b) for b in rng.integers(0, 2, controls))
            ins.append(Instruction(G.Control(G.Unitary(m), bits), q))
            psi = O.np_apply(psi, n, O.np_controlled(m, controls, bits), q[:controls] + q[controls:][::-1])
        else:
            step(Instruction(G.Unitary(m), q))
        step(Instruction(G.Rx(float(rng.uniform(-3, 3))), [int(rng.integers(n))]))
    return ins, psi


@pytest.mark.parametrize("k", [3, 4, 5, 6])
@pytest.mark.parametrize("n,controls", [(6, 0), (7, 1), (11, 0), (13, 2), (16, 0)])
@pytest.mark.parametrize("mode", [0, 1])
def test_dense_k_vs_oracle(k, n, controls, mode):
    if k + controls > n:
        pytest.skip("not enough qubits")
    rng = np.random.default_rng(1000 * k + 10 * n + controls)
    ins, want = _circuit(n, k, rng, controls)
    _say(f"dense k={k} n={n} controls={controls} mode={mode}")
    sim = StateVectorSimulator(n, 0)
    sim.evolve(ins, mode=mode)
    np.testing.assert_allclose(sim.state_vector(), want, atol=TOL)


@pytest.mark.parametrize("k", [4, 5, 6])
def test_dense_k_apply_gate_and_dm(k):
    rng = np.random.default_rng(77 + k)
    n = 7
    u = random_unitary(k, rng)
    t = [int(x) for x in rng.choice(n, k, replace=False)]
    pre = [Instruction(G.H(), [q]) for q in range(n)] + [Instruction(G.Ry(0.3 * q + 0.2), [q]) for q in range(n)]
    _say(f"dense k={k} apply_gate / density matrix")
    sim = StateVectorSimulator(n, 0)
    sim.evolve(pre)
    sim.apply_gate(G.Unitary(u), *t)
    np.testing.assert_allclose(sim.state_vector(), O.evolve_sv(n, pre + [Instruction(G.Unitary(u), t)]), atol=TOL)
    # density matrix (5 or 6 qubits): the gate on the column bits, its
    # conjugate on the row bits, each a k-bit dense pass over 2^(2m - k) groups
    m = max(5, k)
    t2 = [int(x) for x in rng.choice(m, k, replace=False)]
    u2 = random_unitary(k, rng)
    pre2 = [Instruction(G.H(), [q]) for q in range(m)] + [Instruction(G.Rz(0.4 * q + 0.1), [q]) for q in range(m)]
    for mode in (0, 1):
        dm = DensityMatrixSimulator(m, 0)
        dm.evolve(pre2 + [Instruction(G.Unitary(u2), t2)], mode=mode)
        ref = O.evolve_dm(m, pre2 + [Instruction(G.Unitary(u2), t2)], "reference")
        np.testing.assert_allclose(dm.density_matrix().ravel(order="F"), ref, atol=TOL)
