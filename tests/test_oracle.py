"""The oracle (oracle/refcpu.c + oracle/oracle.py) pinned against the
reference's own golden values, closed forms, published known-answer
vectors and numpy brute force.  CPU only."""
import math

import numpy as np
import pytest

import oracle as O
from bsim import gates as G
from bsim import noises as Nz
from bsim.simulators import Instruction
from bsim.gates import kernel_call
from bsim import workloads as W

from circuits import fixture_ops, grcs16, random_circuit, random_noise, random_unitary
from conftest import GOLDEN

# the reference tests compare 8-digit literals with isapprox (rtol ~1.5e-8)
GOLD_ATOL = 6e-8


def _cx(rows):
    return np.array([complex(a, b) for a, b in rows])


def test_sv_simple_circuits(fixtures):
    for case in fixtures["sv_simple"]:
        psi = O.evolve_sv(case["n"], fixture_ops(case))
        np.testing.assert_allclose(psi, _cx(case["state"]), atol=GOLD_ATOL, err_msg=case["cite"])
        np.testing.assert_allclose(O.sv_probabilities(psi, case["n"]), case["probs"], atol=GOLD_ATOL,
                                   err_msg=case["cite"])


def _dm_expected(case):
    n = case["n"]
    N = 2 ** n
    if "dm" in case:
        return np.array(case["dm"], dtype=np.complex128)
    m = np.zeros((N, N), dtype=np.complex128)
    m[case["dm_diag_one"], case["dm_diag_one"]] = 1
    return m


def test_dm_simple_circuits(fixtures):
    for case in fixtures["dm_simple"]:
        n = case["n"]
        N = 2 ** n
        rho = O.evolve_dm(n, fixture_ops(case)).reshape((N, N), order="F")
        np.testing.assert_allclose(rho, _dm_expected(case), atol=GOLD_ATOL, err_msg=case["cite"])
        probs = O.dm_probabilities(rho.ravel(order="F").copy(), n)
        exp = case.get("probs")
        if exp is None:
            exp = np.zeros(N)
            exp[case["probs_one"]] = 1
        np.testing.assert_allclose(probs, exp, atol=GOLD_ATOL, err_msg=case["cite"])


def test_dm_two_qubit_pauli_channel_equals_kraus():
    # test/test_dm_simulator.jl:23-31,78-90
    xx = np.kron([[0, 1], [1, 0]], [[0, 1], [1, 0]])
    yz = np.kron([[0, -1j], [1j, 0]], [[1, 0], [0, -1]])
    kr = Nz.Kraus([math.sqrt(0.7) * np.eye(4), math.sqrt(0.1) * xx, math.sqrt(0.2) * yz])
    base = [Instruction(G.X(), [0]), Instruction(G.X(), [1])]
    a = O.evolve_dm(2, base + [Instruction(kr, [0, 1])])
    b = O.evolve_dm(2, base + [Instruction(Nz.TwoQubitPauliChannel({"XX": 0.1, "YZ": 0.2}), [0, 1])])
    np.testing.assert_allclose(a, b, atol=1e-15)


@pytest.mark.parametrize("n", [4, 9, 14, 20])
def test_qft_uniform(n):
    # test/test_sv_simulator.jl:360-384
    ins = []
    for t in range(n):
        ang = math.pi / 2
        ins.append(Instruction(G.H(), [t]))
        for c in range(t + 1, n):
            ins.append(Instruction(G.CPhaseShift(ang), [c, t]))
            ang /= 2
    p = O.sv_probabilities(O.evolve_sv(n, ins), n)
    np.testing.assert_allclose(p, np.full(2 ** n, 2.0 ** -n), rtol=1e-9)


@pytest.mark.parametrize("n", [2, 6])
def test_dm_qft_uniform(n):
    # test/test_dm_simulator.jl:345-369
    ins = []
    for t in range(n):
        ang = math.pi / 2
        ins.append(Instruction(G.H(), [t]))
        for c in range(t + 1, n):
            ins.append(Instruction(G.CPhaseShift(ang), [c, t]))
            ang /= 2
    p = O.dm_probabilities(O.evolve_dm(n, ins), n)
    np.testing.assert_allclose(p, np.full(2 ** n, 2.0 ** -n), rtol=1e-9)


def test_grcs16_first_probability(fixtures):
    # test/test_openqasm3.jl:986-992
    ins = grcs16(GOLDEN / "grcs_16.qasm")
    p = O.sv_probabilities(O.evolve_sv(16, ins), 16)
    assert abs(p[0] - fixtures["grcs16"]["first_prob"]) <= fixtures["grcs16"]["atol"]
    assert abs(p.sum() - 1) < 1e-12


def test_ry_qft_closed_form():
    """configs[0] circuit == sqrt(N) ifft of the RY product state."""
    ins, th = W.ry_qft(12, seed=0)
    psi = O.evolve_sv(12, ins)
    np.testing.assert_allclose(psi, O.qft_state(th), atol=1e-13)


# ---- refcpu vs numpy brute force (index arithmetic, conventions) ----------

def _np_call(psi, n, call):
    """Brute-force application of a KernelCall with the reference's
    conventions written out independently."""
    m = np.asarray(call.matrix, dtype=np.complex128)
    if call.diagonal:
        m = np.diag(m)
    t = list(call.targets)
    c = list(call.controls)
    if call.double_excitation is not None:
        return O.np_apply(psi, n, m, list(reversed(t)))
    # matrix bit b <-> t[b]  (bit 0 = LSB)  ==> kron order is reversed(t)
    if c:
        full = O.np_controlled(m, len(c), call.cvals)
        return O.np_apply(psi, n, full, c + list(reversed(t)))
    return O.np_apply(psi, n, m, list(reversed(t)))


@pytest.mark.parametrize("seed", range(6))
def test_refcpu_matches_bruteforce(seed):
    rng = np.random.default_rng(seed)
    n = 6
    ins = random_circuit(n, 40, rng)
    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi[0] = 1
    ref = psi.copy()
    for i in ins:
        call = kernel_call(i.operator, i.target)
        if len(call.controls) >= 2 and len(call.targets) == 2:
            continue
        O.apply_call(psi, n, call)
        ref = _np_call(ref, n, call)
    np.testing.assert_allclose(psi, ref, atol=1e-12)


def test_two_qubit_kernel_t1_is_lsb():
    # test/test_sv_simulator.jl:94-99 pins the convention: CPhaseShift01 on
    # [0,1] after X(1) phases index 1
    psi = O.evolve_sv(2, [Instruction(G.X(), [1]), Instruction(G.CPhaseShift01(0.15), [0, 1])])
    assert abs(psi[1] - np.exp(0.15j)) < 1e-15


def test_kq_fallback_ignores_target_order():
    rng = np.random.default_rng(1)
    u = random_unitary(3, rng)
    a = O.evolve_sv(4, [Instruction(G.H(), [q]) for q in range(4)] + [Instruction(G.Unitary(u), [0, 2, 3])])
    b = O.evolve_sv(4, [Instruction(G.H(), [q]) for q in range(4)] + [Instruction(G.Unitary(u), [3, 0, 2])])
    np.testing.assert_allclose(a, b, atol=0)


def _np_dm(n, ops, mode):
    """Physical rho (numpy) for gates and noise; mode 'reference' restates
    the block-diagonal Kraus update independently of refcpu."""
    N = 2 ** n
    S = np.zeros((N, N), dtype=np.complex128)  # stored = rho^T
    S[0, 0] = 1
    for ins in ops:
        op = ins.operator
        if isinstance(op, G.Gate):
            call = kernel_call(op, ins.target)
            U = np.eye(N, dtype=np.complex128)
            for col in range(N):
                U[:, col] = _np_call(np.eye(N, dtype=np.complex128)[:, col].copy(), n, call)
            S = np.conj(U) @ S @ U.T
        else:
            ks = op.kraus_rep()
            t = list(ins.target)
            k = len(t)
            # channel matrix bit -> qubit, as refcpu: k=1; k=2 t1 MSB; k>2 sorted
            kron_q = t if k <= 2 else sorted(t)
            full = []
            for K in ks:
                F = np.eye(N, dtype=np.complex128)
                for col in range(N):
                    F[:, col] = O.np_apply(np.eye(N, dtype=np.complex128)[:, col].copy(), n, K, kron_q)
                full.append(F)
            if mode == "physical" or isinstance(op, Nz.PhaseFlip):
                # PhaseFlip has its own elementwise kernel over all 4^n entries
                # (src/noise_kernels.jl:7-23), so it is never block-diagonal
                S = sum(np.conj(F) @ S @ F.T for F in full)
            else:
                new = S.copy()
                rest = [q for q in range(n) if q not in t]
                for r in range(N):
                    for c in range(N):
                        if all(((r >> (n - 1 - q)) & 1) == ((c >> (n - 1 - q)) & 1) for q in rest):
                            acc = 0
                            for F in full:  # K B K^dagger restricted to the block
                                blk_r = [x for x in range(N) if all(((x >> (n - 1 - q)) & 1) == ((r >> (n - 1 - q)) & 1) for q in rest)]
                                acc += sum(F[r, a] * S[a, b] * np.conj(F[c, b]) for a in blk_r for b in blk_r)
                            new[r, c] = acc
                S = new
    return S


@pytest.mark.parametrize("mode", ["reference", "physical"])
@pytest.mark.parametrize("seed", range(3))
def test_refcpu_dm_matches_bruteforce(mode, seed):
    rng = np.random.default_rng(100 + seed)
    n = 3
    ops = []
    for _ in range(8):
        ops += random_circuit(n, 1, rng, kinds=("1q", "2q"))
        ops.append(random_noise(n, rng))
    got = O.evolve_dm(n, ops, mode=mode).reshape((2 ** n, 2 ** n), order="F")
    exp = _np_dm(n, ops, mode)
    np.testing.assert_allclose(got, exp, atol=1e-12)


def test_blockdiag_differs_from_physical():
    """The reference's Kraus kernel touches only block-diagonal blocks
    (src/noise_kernels.jl:111-131): H(x)H, Depolarizing(0.3) on q0, H(x)H."""
    ops = [Instruction(G.H(), [0]), Instruction(G.H(), [1]), Instruction(Nz.Depolarizing(0.3), [0]),
           Instruction(G.H(), [0]), Instruction(G.H(), [1])]
    ref = O.dm_probabilities(O.evolve_dm(2, ops, "reference"), 2)
    phys = O.dm_probabilities(O.evolve_dm(2, ops, "physical"), 2)
    np.testing.assert_allclose(ref, [0.9, -0.1, 0.1, 0.1], atol=1e-12)
    np.testing.assert_allclose(phys, [0.8, 0.0, 0.2, 0.0], atol=1e-12)


def test_expectations_match_bruteforce():
    rng = np.random.default_rng(7)
    n = 5
    psi = O.evolve_sv(n, random_circuit(n, 30, rng))
    for word in ["ZIIII", "IXIYI", "ZZZZZ", "XYZXY"]:
        full = np.ones((1, 1))
        for ch in word:
            full = np.kron(full, {"I": np.eye(2), "X": O._PAULI["X"], "Y": O._PAULI["Y"], "Z": O._PAULI["Z"]}[ch])
        assert abs(O.sv_expect_pauli(psi, n, word) - np.real(np.vdot(psi, full @ psi))) < 1e-12
    h = random_unitary(2, rng)
    h = h + h.conj().T
    # 2q Hermitian: t1 = MSB (kron order of the given targets)
    exp = np.real(np.vdot(psi, O.np_apply(psi, n, h, [3, 1])))
    assert abs(O.sv_expect_matrix(psi, n, h, [3, 1]) - exp) < 1e-12


def test_marginal_matches_numpy():
    rng = np.random.default_rng(3)
    n = 6
    p = O.sv_probabilities(O.evolve_sv(n, random_circuit(n, 30, rng)), n)
    t = [4, 1]
    m = O.marginal(p, n, t)
    exp = p.reshape([2] * n).sum(axis=tuple(q for q in range(n) if q not in t)).reshape(-1)
    np.testing.assert_allclose(m, exp, atol=1e-15)  # sorted target order (the reference ignores order)


# ---- sampling checkers -------------------------------------------------------

def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32_10 (kat_vectors)
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF, 0xFFFFFFFF]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_splitmix64_known_answer():
    # SplitMix64 seeded with 0: first outputs 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4
    d = O.splitmix_doubles(0, 2)
    assert d[0] == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    assert d[1] == (0x6E789E6AA1B965F4 >> 11) * 2.0 ** -53
    np.testing.assert_array_equal(W.splitmix64_doubles(12345, 64), O.splitmix_doubles(12345, 64))


def _probs(seed, n):
    rng = np.random.default_rng(seed)
    return O.sv_probabilities(O.evolve_sv(n, random_circuit(n, 40, rng)), n)


@pytest.mark.parametrize("n", [1, 3, 8, 12])
def test_alias_reference_samples_p(n):
    p = _probs(n, n)
    ap, al = O.alias_reference(p, p.sum())
    np.testing.assert_allclose(O.alias_distribution(np.minimum(ap, 1.0), al), p, atol=1e-12)


@pytest.mark.parametrize("n", [1, 3, 8, 12, 16])
def test_alias_exact_integer_invariant(n):
    p = _probs(n + 50, n)
    ap, al = O.alias_exact(p, float(p.sum()))
    C = 2.0 ** 32
    own = np.round(ap * C)
    assert np.all(own == ap * C) and np.all(ap <= 1.0) and np.all((ap > 0) | (p == 0))
    # mass of item k = own_k + sum over buckets b aliasing to k of (C - own_b):
    # exactly the integer weight floor(p_k N C / wsum) (+ remainder on the max)
    mass = own.copy()
    np.add.at(mass, al, C - own)
    w = np.floor(p * (2 ** n * C / p.sum()))
    w[np.argmax(w)] += 2 ** n * C - w.sum()
    np.testing.assert_array_equal(mass, w)


def test_alias_exact_special_cases():
    for p in [np.full(8, 1 / 8), np.eye(1, 16, 5).ravel(), np.array([0.5, 0.5, 0, 0]), np.array([1.0])]:
        ap, al = O.alias_exact(p, float(p.sum()))
        np.testing.assert_allclose(O.alias_distribution(ap, al), p, atol=1e-9)


@pytest.mark.parametrize("n", [1, 4, 10])
def test_direct_sample_is_cdf_search(n):
    """StatsBase.sample's linear scan (the reference's n >= 30 sampler) is
    the first index whose running sum reaches u * sum(w): the same index as a
    binary search of numpy's sequential cumsum, clamped to the last entry."""
    p = _probs(n + 7, n)
    rng = np.random.default_rng(n)
    u = np.concatenate([rng.random(2000), [0.0, 1.0 - 2 ** -53, 0.5]])
    got = O.direct_samples(p, float(p.sum()), u)
    exp = np.minimum(np.searchsorted(np.cumsum(p), u * p.sum(), side="left"), p.size - 1)
    np.testing.assert_array_equal(got, exp)


def test_direct_sample_hand_cases():
    w = np.array([0.0, 0.25, 0.0, 0.5, 0.25])
    u = np.array([0.0, 0.1, 0.25, 0.26, 0.74, 0.75, 0.76, 0.999])
    # t = u: index 0 only at t = 0 (cw = 0 is not < 0); zero weights are skipped otherwise
    np.testing.assert_array_equal(O.direct_samples(w, 1.0, u), [0, 1, 1, 3, 3, 3, 4, 4])


def test_shot_rule():
    ap = np.array([1.0, 0.25, 0.5, 1.0])
    al = np.array([0, 3, 0, 3])
    idx = np.array([1, 1, 2, 2, 0])
    u = np.array([0.1, 0.3, 0.49, 0.5, 0.99])
    np.testing.assert_array_equal(O.lookup(ap, al, idx, u), [1, 3, 2, 0, 0])


def test_cdf_shots_special_cases():
    """Inverse-CDF shots of the batch path: a basis state always returns its
    index, and the uniform superposition (every weight exactly 2^32) returns
    the top n bits of (x1:x0) -- the alias path's Philox index draw."""
    n = 6
    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi[37] = 1j
    assert (O.cdf_shots(psi, n, 500, 3) == 37).all()
    uni = np.full(2 ** n, 2 ** (-n / 2), dtype=np.complex128)
    idx, _ = O.shot_draws(11, n, 2000)
    np.testing.assert_array_equal(O.cdf_shots(uni, n, 2000, 11), idx)


def test_cdf_shots_frequencies():
    """Empirical frequencies of 400k shots of a random 4-qubit state within
    5 sigma of |psi_i|^2, and zero-probability states never drawn."""
    rng = np.random.default_rng(2)
    psi = rng.normal(size=16) + 1j * rng.normal(size=16)
    psi[[3, 9]] = 0
    psi /= np.linalg.norm(psi)
    p = np.abs(psi) ** 2
    s = O.cdf_shots(psi, 4, 400000, 5)
    f = np.bincount(s, minlength=16) / len(s)
    assert f[3] == 0 and f[9] == 0
    assert (np.abs(f - p) < 5 * np.sqrt(p * (1 - p) / len(s)) + 1e-12).all()


# ---- reduced density matrices (partial_trace) --------------------------------

def _golden_dm():
    import json
    d = json.loads((GOLDEN / "density_matrix_results.json").read_text())
    cm = lambda rows: np.array([[complex(*x) for x in r] for r in rows])
    k = np.arange(16)
    psi = np.sqrt(k / 120.0) * np.where(k < 8, 1.0, 1j)
    return d, psi, cm(d["density_matrix_3"]), cm(d["density_matrix_0_2_1"])


def test_partial_trace_golden():
    """DensityMatrix(3) and DensityMatrix([0, 2, 1]) of the reference's
    test state (test/test_result_types.jl:44-60, rtol 1e-6); the second one
    takes the unsorted-target reindexing."""
    d, psi, m3, m021 = _golden_dm()
    np.testing.assert_allclose(O.partial_trace(psi, 4, [3], dm=False), m3, rtol=d["rtol"], atol=1e-8)
    np.testing.assert_allclose(O.partial_trace(psi, 4, [0, 2, 1], dm=False), m021, rtol=d["rtol"], atol=1e-8)
    full = np.outer(psi, np.conj(psi))
    # DensityMatrix(reverse(0:3)) == kron(adjoint(sv), sv) (test/test_result_types.jl:212-219)
    np.testing.assert_allclose(O.permute_density_matrix(full, 4, [3, 2, 1, 0]), full, atol=1e-15)


@pytest.mark.parametrize("n", [3, 4, 5])
def test_partial_trace_sv_equals_dm(n):
    """test/test_sv_simulator.jl:654-662: the state-vector and density-matrix
    traces agree for every sorted output set; plus closed forms: the norm for
    no outputs, the full projector for all of them, and a product state's
    reduced matrix is its factor."""
    from itertools import combinations

    rng = np.random.default_rng(n)
    psi = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    psi /= np.linalg.norm(psi)
    full = np.outer(psi, np.conj(psi))
    for r in range(1, n + 1):
        for q in combinations(range(n), r):
            np.testing.assert_allclose(O.partial_trace(psi, n, list(q), dm=False),
                                       O.partial_trace(full, n, list(q), dm=True), atol=1e-13)
    assert abs(O.partial_trace(psi, n, [], dm=False) - 1.0) < 1e-13
    a = rng.normal(size=2) + 1j * rng.normal(size=2)
    a /= np.linalg.norm(a)
    b = rng.normal(size=2 ** (n - 1)) + 1j * rng.normal(size=2 ** (n - 1))
    b /= np.linalg.norm(b)
    prod = np.kron(a, b)  # qubit 0 is the factor a
    np.testing.assert_allclose(O.partial_trace(prod, n, [0], dm=False), np.outer(a, np.conj(a)), atol=1e-13)


@pytest.mark.parametrize("seed", range(3))
def test_physical_superop_form_equals_block_form(seed):
    """The superoperator form of a physical Kraus channel (used for the
    14-qubit checks) equals the block-by-block form, channels of 1, 2 and 3
    qubits on unsorted targets."""
    rng = np.random.default_rng(300 + seed)
    n = 4
    ops = []
    for _ in range(6):
        ops += random_circuit(n, 2, rng, kinds=("1q", "2q"))
        ops.append(random_noise(n, rng))
    ops.append(Instruction(Nz.Kraus([random_unitary(3, rng) * math.sqrt(0.5), random_unitary(3, rng) * math.sqrt(0.5)]),
                           [3, 0, 2]))
    ops.append(Instruction(Nz.TwoQubitDepolarizing(0.1), [2, 0]))
    a = O.evolve_dm(n, ops, mode="physical")
    b = O.evolve_dm(n, ops, mode="physical", physical_as_superop=True)
    np.testing.assert_allclose(b, a, atol=1e-13)
