"""The five benchmark workloads at their full stated sizes on the GPU,
against the oracle (or, where the oracle cannot run the whole workload in
seconds, against the GPU's plain per-op path plus an oracle sample):

  C2  30-qubit brick-wall, depth 20 (16 GiB state): the oracle's own 30-qubit
      evolve (the reference's per-gate loops on the host cores); <Z_i> for
      every qubit and every amplitude within 1e-10;
  C3  14-qubit noisy density matrix (4 GiB), depth 10, both Kraus semantics:
      the whole stored matrix, the diagonal probabilities and <Z_i Z_i+1>
      within 1e-10 of the oracle;
  C4  24- and 28-qubit hardware-efficient ansatz, 8 layers: <H> against the
      oracle (24q), the fused adjoint gradient against the per-op one and
      against parameter shift through the plain path on a parameter sample,
      within 1e-10; the 12-qubit noisy density-matrix variant against the
      oracle's parameter shift on a sample;
  C5  4096 uneven brick-wall circuits and 256 parameter sets of the
      16-qubit ansatz with their adjoint gradients in one persistent launch,
      against one launch per circuit (<Z_q> within 1e-10) and the serial
      adjoint gradient (within 1e-10), and a sample of circuits against the
      oracle (inverse-CDF shots bit-exact).

Each step prints what it runs, so a long step is never silent for long.
"""
import math
import sys
import time

import numpy as np
import pytest

import oracle as O
from bsim import _lib as L
from bsim import adjoint as A
from bsim import workloads as W
from bsim.batch import run_batch
from bsim.gates import kernel_call
from bsim.simulators import DensityMatrixSimulator, StateVectorSimulator

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _say(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)
    sys.stdout.flush()


def _oracle_sv(n, ins, label):
    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi[0] = 1
    t = time.time()
    for k, i in enumerate(ins):
        O.apply_call(psi, n, kernel_call(i.operator, i.target))
        if k % 100 == 99:
            _say(f"{label}: oracle gate {k + 1}/{len(ins)} ({time.time() - t:.0f} s)")
    return psi


def _z_from_probs(p, n):
    """<Z_q> for every qubit from the probabilities (qubit q = bit n-1-q):
    marginals of the top and bottom halves of the index, two passes."""
    h = n // 2
    m = p.reshape(2 ** h, 2 ** (n - h))
    top = m.sum(axis=1)      # qubits 0..h-1
    bot = m.sum(axis=0)      # qubits h..n-1
    z = np.empty(n)
    for q in range(n):
        if q < h:
            v = top.reshape(2 ** q, 2, -1)
        else:
            v = bot.reshape(2 ** (q - h), 2, -1)
        z[q] = v[:, 0, :].sum() - v[:, 1, :].sum()
    return z


def _max_diff_chunked(a, b, chunk=1 << 26):
    m = 0.0
    for s in range(0, a.size, chunk):
        m = max(m, float(np.max(np.abs(a[s:s + chunk] - b[s:s + chunk]))))
    return m


@pytest.mark.timeout(900)
def test_c2_brickwall_30q_vs_oracle():
    n, depth = 30, 20
    ins = W.brickwall(n, depth, seed=0)
    _say(f"C2: GPU fused evolve, {n} qubits, {len(ins)} gates")
    sim = StateVectorSimulator(n, 0)
    sim.evolve(ins)
    z_gpu = sim.expect_z_all()
    got = sim.state_vector()
    del sim
    exp = _oracle_sv(n, ins, "C2")
    _say("C2: comparing amplitudes")
    err = _max_diff_chunked(got, exp)
    del got
    assert err < TOL, f"max |amplitude difference| {err}"
    p = O.sv_probabilities(exp, n)
    del exp
    z = _z_from_probs(p, n)
    _say(f"C2: max |amp diff| {err:.2e}, max |<Z> diff| {np.max(np.abs(z - z_gpu)):.2e}")
    np.testing.assert_allclose(z_gpu, z, atol=TOL)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("mode", ["reference", "physical"])
def test_c3_noisy_dm_14q_vs_oracle(mode):
    n = 14
    ins = W.brickwall(n, 10, seed=0, noise=True)
    _say(f"C3 {mode}: GPU fused evolve, {n} qubits, {len(ins)} ops")
    sim = DensityMatrixSimulator(n, 0, noise_mode=mode)
    sim.evolve(ins)
    words = ["I" * i + "ZZ" + "I" * (n - 2 - i) for i in range(n - 1)]
    zz = sim.expect_pauli_words(words)
    probs = sim.probabilities()
    got = sim.density_matrix().ravel(order="F")
    del sim
    _say(f"C3 {mode}: oracle evolve")
    t = time.time()
    exp = O.evolve_dm(n, ins, mode=mode, physical_as_superop=True)
    _say(f"C3 {mode}: oracle done in {time.time() - t:.0f} s, comparing")
    err = _max_diff_chunked(got, exp)
    del got
    assert err < TOL, f"max |rho difference| {err}"
    pe = O.dm_probabilities(exp, n)
    np.testing.assert_allclose(probs, pe, atol=TOL)
    # <Z_i Z_i+1> is diagonal: sum_k p_k (-1)^(b_i + b_i+1)
    idx = np.arange(2 ** n)
    for i, w in enumerate(words):
        s = 1.0 - 2.0 * (((idx >> (n - 1 - i)) ^ (idx >> (n - 2 - i))) & 1)
        assert abs(zz[i] - float(np.dot(pe, s))) < TOL
    # the channels keep the trace in both semantics (the diagonal lies in the
    # blocks the reference updates)
    assert abs(pe.sum() - 1.0) < 1e-9


def _ps_sample(sim, ins, pmap, th, words, coeffs, sample):
    """Parameter shift through the plain per-op path on a parameter sample."""
    out = []
    for p in sample:
        vals = []
        for s in (math.pi / 2, -math.pi / 2):
            t = th.copy()
            t[p] += s
            sim.reinit(sim.qubit_count, 0)
            sim.evolve(A.with_params(ins, pmap, t), mode=1)
            vals.append(A.expectation_pauli_sum(sim, words, coeffs))
        out.append(0.5 * (vals[0] - vals[1]))
    return np.array(out)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("n", [24, 28])
def test_c4_hea_adjoint(n):
    ins, pmap, th = W.hea(n, 8, seed=0)
    words, coeffs = W.tfim(n)
    sim = StateVectorSimulator(n, 0)
    _say(f"C4: {n}-qubit ansatz, {len(th)} parameters: fused adjoint")
    ev, g = A.adjoint_gradient(sim, ins, pmap, words, coeffs)
    _say(f"C4 {n}q: per-op adjoint")
    ev2, g2 = A.adjoint_gradient(sim, ins, pmap, words, coeffs, fused=False)
    assert abs(ev - ev2) < TOL
    np.testing.assert_allclose(g, g2, atol=TOL)
    rng = np.random.default_rng(n)
    sample = sorted(set(int(x) for x in rng.choice(len(th), 12, replace=False)) | {0, len(th) - 1})
    _say(f"C4 {n}q: parameter shift on {len(sample)} parameters")
    ps = _ps_sample(sim, ins, pmap, th, words, coeffs, sample)
    np.testing.assert_allclose(g[sample], ps, atol=TOL)
    if n == 24:
        _say("C4 24q: oracle <H>")
        psi = _oracle_sv(n, ins, "C4")
        e = sum(c * O.sv_expect_pauli(psi, n, w) for w, c in zip(words, coeffs))
        assert abs(ev - e) < TOL


@pytest.mark.timeout(600)
def test_c4_noisy_dm_adjoint_12q():
    n = 12
    ins, pmap, th = W.hea(n, 8, seed=0, noise_p=0.005)
    words, coeffs = W.tfim(n)
    sim = DensityMatrixSimulator(n, 0)
    _say(f"C4 noisy DM: {n} qubits, {len(ins)} ops, {len(th)} parameters")
    ev, g = A.adjoint_gradient(sim, ins, pmap, words, coeffs)
    rng = np.random.default_rng(12)
    sample = sorted(set(int(x) for x in rng.choice(len(th), 3, replace=False)) | {len(th) - 1})

    def energy(t):
        rho = O.evolve_dm(n, A.with_params(ins, pmap, t), mode="reference")
        return sum(c * O.dm_expect_pauli(rho, n, w) for w, c in zip(words, coeffs))

    _say("C4 noisy DM: oracle <H>")
    assert abs(ev - energy(th)) < TOL
    for p in sample:
        _say(f"C4 noisy DM: oracle parameter shift, parameter {p}")
        tp, tm = th.copy(), th.copy()
        tp[p] += math.pi / 2
        tm[p] -= math.pi / 2
        assert abs(g[p] - 0.5 * (energy(tp) - energy(tm))) < TOL


@pytest.mark.timeout(600)
def test_c5_batch_4096():
    """The 4096 circuits and 256 parameter sets of the 16-qubit ansatz (each
    with its adjoint gradient) in one persistent launch."""
    from bsim.batch import GradientTask

    count, shots, nsets = 4096, 1000, 256
    spec = W.batch_tasks(count, seed=0)
    circuits = [(n, W.brickwall_ops(n, d, s)) for n, d, s in spec]
    seeds = [s for _, _, s in spec]
    words, coeffs = W.tfim(16)
    sets = [W.hea(16, 8, seed=k) for k in range(nsets)]
    grads = [GradientTask(16, ins, pmap, words, coeffs) for ins, pmap, _ in sets]
    _say(f"C5: {count} circuits and {nsets} gradient tasks in one persistent launch")
    allres = run_batch(circuits + grads, [shots] * count + [0] * nsets, seeds + [0] * nsets)
    res, gres = allres[:count], allres[count:]
    _say("C5: serial adjoint gradients")
    gsim = StateVectorSimulator(16, 0)
    for k, (ins, pmap, _) in enumerate(sets):
        ev, g = A.adjoint_gradient(gsim, ins, pmap, words, coeffs)
        assert abs(gres[k].expval - ev) < TOL, k
        np.testing.assert_allclose(gres[k].grad, g, atol=TOL, err_msg=f"parameter set {k}")
    del gsim
    _say("C5: one launch per circuit")
    sims = {}
    zmax = 0.0
    for i, (n, ops) in enumerate(circuits):
        sim = sims.get(n) or sims.setdefault(n, StateVectorSimulator(n, 0))
        sim.reinit(n, 0)
        L.call("bsim_evolve", sim.handle, ops.ptr, ops.n, 0)
        zmax = max(zmax, float(np.max(np.abs(sim.expect_z_all() - res[i].expect_z))))
        if i % 1024 == 1023:
            _say(f"C5: per-circuit {i + 1}/{count}")
    del sims
    assert zmax < TOL, zmax
    pick = list(range(0, count, 97))
    _say(f"C5: oracle on {len(pick)} circuits")
    for i in pick:
        n, d, s = spec[i]
        psi = O.evolve_sv(n, W.brickwall(n, d, s))
        assert (O.cdf_shots(psi, n, shots, seeds[i]) == res[i].samples).all(), i
        ze = np.array([O.sv_expect_pauli(psi, n, "I" * q + "Z" + "I" * (n - 1 - q)) for q in range(n)])
        assert np.max(np.abs(ze - res[i].expect_z)) < TOL, i
