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


@pytest.fixture
def grid_cap():
    """Cap the per-gate kernels' grids (bsim_set_max_blocks) so that small
    states walk the grid-stride loops -- and the next-block register
    prefetch of the matrix-core kernels -- that states above about 20
    qubits take with the default grid; restored afterwards."""
    import ctypes

    from bsim import _lib as L

    old = ctypes.c_int()
    L.call("bsim_set_max_blocks", 0, ctypes.byref(old))

    def set_cap(blocks):
        L.call("bsim_set_max_blocks", blocks, None)

    yield set_cap
    L.call("bsim_set_max_blocks", old.value, None)


@pytest.mark.parametrize("k", [3, 4, 5, 6])
@pytest.mark.parametrize("blocks", [1, 3, 7])
@pytest.mark.parametrize("n,controls", [(12, 0), (14, 1), (16, 0), (16, 2)])
def test_dense_k_grid_stride(k, blocks, n, controls, grid_cap):
    """Dense k-target gates with at most `blocks` workgroups per launch: each
    workgroup takes many blocks of 64 groups in turn (odd caps leave a
    partial last round), low / high / scattered targets via the random
    target choice, with and without controls, against the oracle."""
    rng = np.random.default_rng(31 * k + 7 * blocks + n + controls)
    ins, want = _circuit(n, k, rng, controls)
    _say(f"dense k={k} n={n} controls={controls} grid cap {blocks}")
    grid_cap(blocks)
    sim = StateVectorSimulator(n, 0)
    sim.evolve(ins, mode=1)
    got = sim.state_vector()
    grid_cap(0)
    np.testing.assert_allclose(got, want, atol=TOL)


@pytest.mark.parametrize("tg", [[0, 1, 2, 3, 4, 5], [26, 25, 24, 23, 22, 21], [0, 9, 15, 22, 26, 5]])
def test_dense_k_large_state_roundtrip(tg):
    """At 27 qubits (the default grid strides): U then U^dagger for k = 4, 5,
    6 on low, high and scattered targets, with and without a control, gives
    back the state within 1e-12, and U alone moves it (the norm is kept)."""
    n = 27
    rng = np.random.default_rng(len(tg) + tg[1])
    sim = StateVectorSimulator(n, 0)
    prep = [Instruction(G.Ry(float(a)), [q]) for q, a in enumerate(rng.uniform(0.1, 3.0, n))]
    sim.evolve(prep)
    psi0 = sim.state_vector()
    for k in (4, 5, 6):
        for controls in (0, 1):
            q = tg[:k]
            u = random_unitary(k, rng)
            if controls:
                c = [x for x in (13, 17, 19) if x not in q][0]
                fwd = Instruction(G.Control(G.Unitary(u), (1,)), [c] + q)
                bwd = Instruction(G.Control(G.Unitary(u.conj().T), (1,)), [c] + q)
            else:
                fwd, bwd = Instruction(G.Unitary(u), q), Instruction(G.Unitary(u.conj().T), q)
            _say(f"27q round trip k={k} controls={controls} targets={q}")
            sim.evolve([fwd], mode=1)
            mid = sim.state_vector()
            assert abs(np.vdot(mid, mid).real - 1.0) < 1e-11
            assert np.max(np.abs(mid - psi0)) > 1e-3
            sim.evolve([bwd], mode=1)
            np.testing.assert_allclose(sim.state_vector(), psi0, atol=1e-12)


# ---------------------------------------------------------------------------
# beyond the matrix-core kernels: 7..12 targets (k_dense_big), Kraus
# channels on 6..8 qubits (reference semantics, blocks in device memory)
# and 4..5 qubits (physical semantics, 8- and 10-bit superoperators)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("k", [7, 8])
@pytest.mark.parametrize("n,controls", [(9, 0), (10, 1), (12, 0)])
@pytest.mark.parametrize("mode", [0, 1])
def test_dense_big_vs_oracle(k, n, controls, mode):
    """7- and 8-target Unitary gates (random, unsorted targets; with a
    control: Control of the Unitary) as non-tile steps of a fused evolve and
    per gate, against the oracle's k-qubit fallback at 1e-12."""
    if k + controls > n:
        pytest.skip("not enough qubits")
    rng = np.random.default_rng(500 * k + 10 * n + controls)
    ins, want = _circuit(n, k, rng, controls)
    _say(f"dense big k={k} n={n} controls={controls} mode={mode}")
    sim = StateVectorSimulator(n, 0)
    sim.evolve(ins, mode=mode)
    np.testing.assert_allclose(sim.state_vector(), want, atol=TOL)


@pytest.mark.parametrize("k,n", [(7, 7), (9, 11), (10, 11), (11, 12)])
def test_apply_kq_up_to_11_targets(k, n):
    """bsim_apply_kq (apply_gate of a Unitary) on 7..11 targets, including
    every qubit of the register (one group), against the oracle."""
    rng = np.random.default_rng(900 + k)
    u = random_unitary(k, rng)
    t = [int(x) for x in rng.choice(n, k, replace=False)]
    pre = [Instruction(G.Ry(float(a)), [q]) for q, a in enumerate(rng.uniform(0.1, 3.0, n))]
    _say(f"apply_kq k={k} n={n} targets={t}")
    sim = StateVectorSimulator(n, 0)
    sim.evolve(pre)
    sim.apply_gate(G.Unitary(u), *t)
    np.testing.assert_allclose(sim.state_vector(), O.evolve_sv(n, pre + [Instruction(G.Unitary(u), t)]), atol=TOL)


def test_dense_big_density_matrix():
    """A 7-target Unitary on an 8-qubit density matrix (U on the column bits,
    conj(U) on the row bits: two 7-bit passes over the 16-bit vector)."""
    rng = np.random.default_rng(4242)
    m, k = 8, 7
    t = [int(x) for x in rng.choice(m, k, replace=False)]
    u = random_unitary(k, rng)
    pre = [Instruction(G.H(), [q]) for q in range(m)] + [Instruction(G.Rz(0.4 * q + 0.1), [q]) for q in range(m)]
    ins = pre + [Instruction(G.Unitary(u), t)]
    ref = O.evolve_dm(m, ins, "reference")
    for mode in (0, 1):
        _say(f"dense big density matrix mode={mode}")
        dm = DensityMatrixSimulator(m, 0)
        dm.evolve(ins, mode=mode)
        np.testing.assert_allclose(dm.density_matrix().ravel(order="F"), ref, atol=TOL)


def _random_channel(k, nk, rng):
    """nk Kraus matrices of a random k-qubit channel: the d-row blocks of a
    random isometry (nk d x d), so sum K^dagger K = I."""
    d = 2 ** k
    a = rng.normal(size=(nk * d, d)) + 1j * rng.normal(size=(nk * d, d))
    v, _ = np.linalg.qr(a)  # orthonormal columns
    return [v[i * d:(i + 1) * d, :] for i in range(nk)]


@pytest.mark.parametrize("k,n", [(6, 7), (7, 8), (8, 8)])
def test_kraus_reference_semantics_6_to_8_qubits(k, n):
    """Reference-semantics Kraus channels on 6..8 qubits (block-diagonal
    update, blocks in device memory), in a noisy circuit, against the
    oracle's restatement of src/noise_kernels.jl:170-196."""
    from bsim import noises as Nz

    rng = np.random.default_rng(60 + k)
    t = [int(x) for x in rng.choice(n, k, replace=False)]
    ins = [Instruction(G.H(), [q]) for q in range(n)] + [Instruction(G.Ry(0.3 * q + 0.2), [q]) for q in range(n)]
    ins += [Instruction(Nz.Kraus(_random_channel(k, 2, rng)), t), Instruction(G.CNot(), [0, n - 1]),
            Instruction(Nz.Kraus(_random_channel(k, 2, rng)), list(reversed(t)))]
    ref = O.evolve_dm(n, ins, "reference")
    for mode in (0, 1):
        _say(f"kraus reference k={k} n={n} mode={mode}")
        dm = DensityMatrixSimulator(n, 0, noise_mode="reference")
        dm.evolve(ins, mode=mode)
        np.testing.assert_allclose(dm.density_matrix().ravel(order="F"), ref, atol=TOL)


@pytest.mark.parametrize("k", [4, 5])
def test_kraus_physical_4_and_5_qubits(k):
    """Physical Kraus channels on 4 and 5 qubits (8- and 10-bit
    superoperators through the dense kernels) against the oracle's block
    form of sum_m K rho K^dagger; the trace is kept."""
    from bsim import noises as Nz

    n = 6
    rng = np.random.default_rng(80 + k)
    t = [int(x) for x in rng.choice(n, k, replace=False)]
    ins = [Instruction(G.H(), [q]) for q in range(n)] + [Instruction(G.Rx(0.5 * q + 0.1), [q]) for q in range(n)]
    ins += [Instruction(Nz.Kraus(_random_channel(k, 3, rng)), t), Instruction(G.CZ(), [1, 4])]
    ref = O.evolve_dm(n, ins, "physical")
    for mode in (0, 1):
        _say(f"kraus physical k={k} mode={mode}")
        dm = DensityMatrixSimulator(n, 0, noise_mode="physical")
        dm.evolve(ins, mode=mode)
        got = dm.density_matrix()
        np.testing.assert_allclose(got.ravel(order="F"), ref, atol=TOL)
        assert abs(np.trace(got).real - 1.0) < 1e-12


@pytest.mark.parametrize("n", [22, 24])
def test_merged_dense_runs_vs_oracle(n):
    """On vectors of 20+ bits the planner multiplies runs of adjacent dense
    ops around 3+-target gates into one op of at most 6 bits (one
    matrix-core launch instead of several passes): the fused evolve against
    the oracle and against the per-op path."""
    rng = np.random.default_rng(90 + n)
    ins = []
    for _ in range(6):
        k = int(rng.integers(3, 5))
        q = [int(x) for x in rng.choice(n, k, replace=False)]
        ins.append(Instruction(G.Unitary(random_unitary(k, rng)), q))
        ins.append(Instruction(G.Rx(float(rng.uniform(-3, 3))), [q[0]]))
        q2 = [q[1]] + [int(x) for x in rng.choice([x for x in range(n) if x not in q], 2, replace=False)]
        ins.append(Instruction(G.Unitary(random_unitary(3, rng)), q2))
        ins.append(Instruction(G.H(), [int(rng.integers(n))]))
    _say(f"merged dense runs n={n}")
    sim = StateVectorSimulator(n, 0)
    sim.evolve(ins, mode=0)
    got = sim.state_vector()
    np.testing.assert_allclose(got, O.evolve_sv(n, ins), atol=1e-10)
    sim.reinit(n, 0)
    sim.evolve(ins, mode=1)
    np.testing.assert_allclose(got, sim.state_vector(), atol=1e-12)


def test_merged_dm_gate_halves_vs_oracle():
    """3-target gates on an 11-qubit density matrix (22-bit vector): each
    gate's column and row halves merge into one 6-bit op."""
    rng = np.random.default_rng(123)
    m = 11
    ins = [Instruction(G.H(), [q]) for q in range(m)]
    for _ in range(4):
        ins.append(Instruction(G.Unitary(random_unitary(3, rng)), [int(x) for x in rng.choice(m, 3, replace=False)]))
        ins.append(Instruction(G.Ry(float(rng.uniform(-2, 2))), [int(rng.integers(m))]))
    ref = O.evolve_dm(m, ins, "reference")
    for mode in (0, 1):
        _say(f"merged dm halves mode={mode}")
        dm = DensityMatrixSimulator(m, 0)
        dm.evolve(ins, mode=mode)
        np.testing.assert_allclose(dm.density_matrix().ravel(order="F"), ref, atol=1e-10)
