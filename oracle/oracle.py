"""Python side of the oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``liboracle.so`` (refcpu.c, the C restatement of the reference's
kernels) and adds numpy checkers that do not depend on it:

* ``np_apply`` / ``np_state``: brute-force state-vector simulation by
  tensor contraction in plain kron order (qubit 0 = most significant bit),
  used to pin refcpu's index arithmetic on small registers;
* ``qft_state``: the closed form of the RY-product + textbook QFT circuit,
  ``sqrt(N) * ifft(a)``.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this module.  The product path (libbsim, bsim) never does.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)

_SIGS = {
    "ref_apply_1q": [_dp, C.c_int, C.c_int, _dp],
    "ref_apply_2q": [_dp, C.c_int, C.c_int, C.c_int, _dp],
    "ref_apply_ctrl_1c1t": [_dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp],
    "ref_apply_ctrl_1c2t": [_dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp],
    "ref_apply_ctrl_2c1t": [_dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp],
    "ref_apply_kq": [_dp, C.c_int, C.c_int, _ip, _dp],
    "ref_apply_kq_sparse": [_dp, C.c_int, C.c_int, _ip, _dp],
    "ref_apply_double_excitation": [_dp, C.c_int, _ip, C.c_double, C.c_double],
    "ref_apply_hermitian_2q": [_dp, C.c_int, C.c_int, C.c_int, _dp],
    "ref_dm_phaseflip": [_dp, C.c_int, C.c_int, C.c_double],
    "ref_dm_kraus_ref": [_dp, C.c_int, C.c_int, _ip, C.c_int, _dp],
    "ref_dm_kraus_physical": [_dp, C.c_int, C.c_int, _ip, C.c_int, _dp],
    "ref_dm_apply_hermitian": [_dp, C.c_int, C.c_int, _ip, _dp],
    "ref_sv_probabilities": [_dp, C.c_int, _dp],
    "ref_dm_probabilities": [_dp, C.c_int, _dp],
    "ref_dot": [_dp, _dp, C.c_int64, _dp],
    "ref_marginal_probability": [_dp, C.c_int, C.c_int, _ip, _dp],
    "ref_make_alias_table": [_dp, C.c_int64, C.c_double, _dp, _i64p, _i64p, _i64p],
    "ref_direct_sample": [_dp, C.c_int64, C.c_double, C.c_int64, _dp, _i64p],
    "ref_alias_lookup": [_dp, _i64p, C.c_int64, _i64p, _dp, _i64p],
    "ref_philox4x32_10": [_u32p, C.c_uint32, C.c_uint32, _u32p],
    "ref_shot_draws": [C.c_uint64, C.c_int, C.c_int64, C.c_int64, _i64p, _dp],
    "ref_alias_exact": [_dp, C.c_int64, C.c_double, _dp, _i32p],
    "ref_splitmix64_doubles": [C.c_uint64, C.c_int64, _dp],
    "ref_cdf_shots": [_dp, C.c_int, C.c_int64, C.c_uint64, _i64p],
}

_lib = None


def build() -> Path:
    r = subprocess.run(["make", "-C", str(HERE), "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = C.CDLL(str(LIB))
        for name, args in _SIGS.items():
            f = getattr(l, name)
            f.argtypes = args
            f.restype = C.c_double if name == "ref_dm_trace_re" else None
        _lib = l
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _cm(m):
    return np.asfortranarray(np.asarray(m, dtype=np.complex128))


# ---------------------------------------------------------------------------
# kernel-level calls (the same description the host gives libbsim)
# ---------------------------------------------------------------------------

def apply_call(vec: np.ndarray, vq: int, call, conj: bool = False, shift: int = 0):
    """Apply one reference kernel (described by a bsim.gates.KernelCall) to
    ``vec`` (length 2^vq), with qubits shifted by ``shift`` (the DM row pass)."""
    L = lib()
    m = np.asarray(call.matrix, dtype=np.complex128)
    if call.diagonal:
        m = np.diag(m)
    if conj:
        m = np.conj(m)
    m = _cm(m)
    t = [x + shift for x in call.targets]
    c = [x + shift for x in call.controls]
    if call.double_excitation is not None:
        orig = np.array(list(reversed(t)), dtype=np.int32)
        a = call.double_excitation
        L.ref_apply_double_excitation(_d(vec), vq, orig.ctypes.data_as(_ip), math.cos(a / 2.0), math.sin(a / 2.0))
        return
    if c:
        if len(c) == 1 and len(t) == 1:
            L.ref_apply_ctrl_1c1t(_d(vec), vq, int(call.cvals[0]), c[0], t[0], _d(m))
        elif len(c) == 1 and len(t) == 2:
            L.ref_apply_ctrl_1c2t(_d(vec), vq, int(call.cvals[0]), c[0], t[0], t[1], _d(m))
        elif len(c) == 2 and len(t) == 1:
            L.ref_apply_ctrl_2c1t(_d(vec), vq, int(call.cvals[0]), int(call.cvals[1]), c[0], c[1], t[0], _d(m))
        else:
            raise NotImplementedError("the reference has no kernel for this control arity")
        return
    if len(t) == 1:
        L.ref_apply_1q(_d(vec), vq, t[0], _d(m))
    elif len(t) == 2:
        L.ref_apply_2q(_d(vec), vq, t[0], t[1], _d(m))
    else:
        ta = np.array(t, dtype=np.int32)
        L.ref_apply_kq(_d(vec), vq, len(t), ta.ctypes.data_as(_ip), _d(m))


def evolve_sv(n: int, instructions, psi: np.ndarray | None = None) -> np.ndarray:
    """evolve! (src/sv_simulator.jl:140-148) on the oracle."""
    from bsim.gates import Gate, kernel_call  # host gate -> kernel description

    if psi is None:
        psi = np.zeros(2 ** n, dtype=np.complex128)
        psi[0] = 1
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    for ins in instructions:
        op = ins.operator
        if not isinstance(op, Gate):
            continue
        apply_call(psi, n, kernel_call(op, ins.target))
    return psi


def _kraus_kron_qubits(targets):
    """Qubits of a Kraus matrix's kron order (first = MSB) as the reference's
    kernels index it: 1 and 2 qubits as given (2q: t1 is the MSB,
    src/noise_kernels.jl:144-155), k > 2 ascending (sorted endian targets,
    :175-186)."""
    t = [int(x) for x in targets]
    return t if len(t) <= 2 else sorted(t)


def physical_channel_superop(ks, targets, n: int):
    """The physical channel S <- sum_m conj(K_m) S K_m^T on the stored matrix
    S = rho^T as one operator on the 4^n vector: S'[r, c] = sum
    conj(K)[r, r'] K[c, c'] S[r', c'], i.e. kron(K, conj(K)) over the
    channel's column bits (vector qubits t) then row bits (vector qubits
    n + t).  Returns (matrix in kron order of the ascending vector qubits,
    those vector qubits): the convention of ref_apply_kq."""
    kq = _kraus_kron_qubits(targets)
    k = len(kq)
    sup = sum(np.kron(np.asarray(K, dtype=np.complex128), np.conj(np.asarray(K, dtype=np.complex128))) for K in ks)
    src = list(kq) + [n + q for q in kq]
    dst = sorted(src)
    perm = [src.index(q) for q in dst]
    t = sup.reshape([2] * (4 * k))
    m = t.transpose(perm + [2 * k + p for p in perm]).reshape(4 ** k, 4 ** k)
    return m, dst


def evolve_dm(n: int, instructions, mode: str = "reference", rho: np.ndarray | None = None,
              physical_as_superop: bool = False) -> np.ndarray:
    """evolve! (src/dm_simulator.jl:170-179) on the oracle.  Returns the
    column-major stored matrix flattened (length 4^n).  ``physical_as_superop``
    (physical mode only) applies each Kraus channel as its superoperator
    on its 2k vector qubits
    instead of block by block, skipping the superoperator's zero entries
    (ref_apply_kq_sparse): the same map, pinned against the block form on
    small registers, and fast enough for 14-qubit matrices."""
    from bsim.gates import Gate, kernel_call
    from bsim.noises import Noise, PhaseFlip

    L = lib()
    if rho is None:
        rho = np.zeros(4 ** n, dtype=np.complex128)
        rho[0] = 1
    rho = np.ascontiguousarray(rho, dtype=np.complex128)
    for ins in instructions:
        op = ins.operator
        if isinstance(op, Gate):
            call = kernel_call(op, ins.target)
            apply_call(rho, 2 * n, call, conj=False, shift=0)   # column bits
            apply_call(rho, 2 * n, call, conj=True, shift=n)    # row bits, conj(U)
        elif isinstance(op, Noise):
            t = np.array(ins.target, dtype=np.int32)
            if isinstance(op, PhaseFlip):
                L.ref_dm_phaseflip(_d(rho), n, int(t[0]), float(op.probability))
                continue
            ks = op.kraus_rep()
            if mode == "physical" and physical_as_superop:
                m, vq = physical_channel_superop(ks, ins.target, n)
                ta = np.array(vq, dtype=np.int32)
                L.ref_apply_kq_sparse(_d(rho), 2 * n, len(vq), ta.ctypes.data_as(_ip), _d(_cm(m)))
                continue
            flat = np.ascontiguousarray(np.concatenate([_cm(k).ravel(order="F") for k in ks]))
            f = L.ref_dm_kraus_physical if mode == "physical" else L.ref_dm_kraus_ref
            f(_d(rho), n, len(t), t.ctypes.data_as(_ip), len(ks), _d(flat))
    return rho


def sv_probabilities(psi: np.ndarray, n: int) -> np.ndarray:
    out = np.empty(2 ** n)
    lib().ref_sv_probabilities(_d(psi), n, _d(out))
    return out


def dm_probabilities(rho: np.ndarray, n: int) -> np.ndarray:
    out = np.empty(2 ** n)
    lib().ref_dm_probabilities(_d(rho), n, _d(out))
    return out


def dot(a: np.ndarray, b: np.ndarray) -> complex:
    out = np.empty(2)
    lib().ref_dot(_d(a), _d(b), a.size, _d(out))
    return complex(out[0], out[1])


_PAULI = {
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def sv_expect_pauli(psi: np.ndarray, n: int, word: str) -> float:
    """real(dot(psi, O psi)) with the Pauli gates applied to a copy
    (src/sv_simulator.jl:240-248, src/result_types.jl:196-221)."""
    L = lib()
    k = psi.copy()
    for q, ch in enumerate(word):
        if ch in _PAULI:
            L.ref_apply_1q(_d(k), n, q, _d(_cm(_PAULI[ch])))
    return dot(psi, k).real


def dm_expect_pauli(rho: np.ndarray, n: int, word: str) -> float:
    """Pauli gates on the column bits of a copy, then real(tr)
    (src/dm_simulator.jl:181-189, :253-260)."""
    L = lib()
    k = rho.copy()
    for q, ch in enumerate(word):
        if ch in _PAULI:
            L.ref_apply_1q(_d(k), 2 * n, q, _d(_cm(_PAULI[ch])))
    N = 2 ** n
    return float(np.real(np.trace(k.reshape((N, N), order="F"))))


def sv_expect_matrix(psi: np.ndarray, n: int, m: np.ndarray, targets) -> float:
    """Hermitian observable on a copy: 1q (:162-178), 2q t1 = MSB
    (:179-198), k > 2 sorted (:200-230)."""
    L = lib()
    k = psi.copy()
    mm = _cm(m)
    t = list(targets)
    if len(t) == 1:
        L.ref_apply_1q(_d(k), n, t[0], _d(mm))
    elif len(t) == 2:
        L.ref_apply_hermitian_2q(_d(k), n, t[0], t[1], _d(mm))
    else:
        ta = np.array(t, dtype=np.int32)
        L.ref_apply_kq(_d(k), n, len(t), ta.ctypes.data_as(_ip), _d(mm))
    return dot(psi, k).real


def dm_expect_matrix(rho: np.ndarray, n: int, m: np.ndarray, targets) -> float:
    L = lib()
    k = rho.copy()
    ta = np.array(list(targets), dtype=np.int32)
    L.ref_dm_apply_hermitian(_d(k), n, len(ta), ta.ctypes.data_as(_ip), _d(_cm(m)))
    N = 2 ** n
    return float(np.real(np.trace(k.reshape((N, N), order="F"))))


_STD = {"X": _PAULI["X"], "Y": _PAULI["Y"], "Z": _PAULI["Z"],
        "H": np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2.0), "I": np.eye(2, dtype=np.complex128)}


def _apply_factor(k: np.ndarray, n: int, factor, targets, dm: bool):
    """apply_observable! of one factor (a letter of a standard observable, or
    a Hermitian matrix) on a copy: standard observables are their gates
    (src/result_types.jl:196-201; on a density matrix the gate acts on the
    column bits of vec(rho), src/dm_simulator.jl:181-189); Hermitian
    matrices go to src/sv_simulator.jl:162-230 (1q; 2q t1 = MSB; k > 2
    sorted) or src/dm_simulator.jl:190-224 (transpose, sorted, every k)."""
    L = lib()
    t = list(targets)
    if isinstance(factor, str):
        L.ref_apply_1q(_d(k), 2 * n if dm else n, t[0], _d(_cm(_STD[factor])))
        return
    mm = _cm(factor)
    if dm:
        ta = np.array(t, dtype=np.int32)
        L.ref_dm_apply_hermitian(_d(k), n, len(ta), ta.ctypes.data_as(_ip), _d(mm))
    elif len(t) == 1:
        L.ref_apply_1q(_d(k), n, t[0], _d(mm))
    elif len(t) == 2:
        L.ref_apply_hermitian_2q(_d(k), n, t[0], t[1], _d(mm))
    else:
        ta = np.array(t, dtype=np.int32)
        L.ref_apply_kq(_d(k), n, len(t), ta.ctypes.data_as(_ip), _d(mm))


def _factor_qubits(f) -> int:
    return 1 if isinstance(f, str) else int(round(math.log2(np.asarray(f).shape[0])))


def expect_factors(state: np.ndarray, n: int, factors, targets, dm: bool) -> float:
    """expectation(sim, TensorProduct(factors) or the single factor, targets...):
    the factors applied in turn to a copy of the state, each on its own
    slice of the targets (src/result_types.jl:203-220), then real(dot(psi,
    copy)) (src/sv_simulator.jl:240-248) or real(tr) (src/dm_simulator.jl:
    253-260).  ``state`` is the state vector or the stored 4^n vector."""
    k = np.ascontiguousarray(state, dtype=np.complex128).copy()
    ix = 0
    for f in factors:
        q = _factor_qubits(f)
        _apply_factor(k, n, f, targets[ix:ix + q], dm)
        ix += q
    if dm:
        N = 2 ** n
        return float(np.real(np.trace(k.reshape((N, N), order="F"))))
    return dot(np.ascontiguousarray(state, dtype=np.complex128), k).real


def expect_op_squared(state: np.ndarray, n: int, factors, targets, dm: bool) -> float:
    """expectation_op_squared (src/result_types.jl:168-194): 1 when every
    factor is a standard observable, else standard factors become I and
    Hermitian ones M*M."""
    if all(isinstance(f, str) for f in factors):
        return 1.0
    sq = ["I" if isinstance(f, str) else np.asarray(f) @ np.asarray(f) for f in factors]
    return expect_factors(state, n, sq, targets, dm)


def calculate_expectation(state, n, factors, targets, dm):
    """calculate(::Expectation) (src/result_types.jl:160-166): one value, or
    one per target when the target count is not the observable's."""
    t = list(targets) if targets else list(range(n))
    nq = sum(_factor_qubits(f) for f in factors)
    if len(t) == nq:
        return expect_factors(state, n, factors, t, dm)
    return [expect_factors(state, n, factors, [q], dm) for q in t]


def calculate_variance(state, n, factors, targets, dm):
    """calculate(::Variance) (src/result_types.jl:223-238)."""
    t = list(targets) if targets else list(range(n))
    nq = sum(_factor_qubits(f) for f in factors)
    if len(t) == nq:
        return expect_op_squared(state, n, factors, t, dm) - expect_factors(state, n, factors, t, dm) ** 2
    return [expect_op_squared(state, n, factors, [q], dm) - expect_factors(state, n, factors, [q], dm) ** 2
            for q in t]


def marginal(probs: np.ndarray, n: int, targets) -> np.ndarray:
    t = np.array(list(targets), dtype=np.int32)
    out = np.empty(2 ** len(t))
    lib().ref_marginal_probability(_d(probs), n, len(t), t.ctypes.data_as(_ip), _d(out))
    return out


# ---------------------------------------------------------------------------
# reduced density matrices (pure Python loops: small registers only)
# ---------------------------------------------------------------------------

def _jl_shl(x: int, s: int) -> int:
    return x << s if s >= 0 else x >> -s


def _pad_bits(ix: int, to_pad) -> int:
    """pad_bit / pad_bits (src/gate_kernels.jl:34-38, :78-84)."""
    for bit in to_pad:
        left = (ix >> bit) << bit
        ix = (left << 1) ^ (ix - left)
    return ix


def _swap_bits(ix: int, qubit_map) -> int:
    """swap_bits (src/dm_simulator.jl:264-278); flip_bit is ix ^ (1 << bit)
    (src/gate_kernels.jl:119-121) with Julia's shift semantics."""
    for in_q, out_q in qubit_map:
        if in_q < out_q:
            in_val = _jl_shl(_jl_shl(1, in_q) & ix, -in_q)
            out_val = _jl_shl(_jl_shl(1, out_q) & ix, -out_q)
            if in_val != out_val:
                ix = ix ^ _jl_shl(1, in_q) ^ _jl_shl(1, out_q)
    return ix


def partial_trace(state: np.ndarray, n: int, output_qubits, dm: bool):
    """partial_trace for a state vector (src/sv_simulator.jl:272-316) or a
    density matrix given as the reference stores it (src/dm_simulator.jl:
    280-323), loop for loop.  ``state``: the 2^n vector, or the 2^n x 2^n
    matrix."""
    from itertools import combinations

    out_q = list(output_qubits)
    if not out_q:
        return np.trace(state) if dm else float(np.sum(np.abs(state) ** 2))
    if len(set(out_q)) == n:
        return state if dm else np.outer(state, np.conj(state))
    qubits = [q for q in range(n) if q not in out_q]
    endian = sorted(n - q - 1 for q in qubits)
    combos = [()] + [c for r in range(1, len(endian) + 1) for c in combinations(endian, r)]
    D = 2 ** (n - len(qubits))
    rho = np.zeros((D, D), dtype=np.complex128)
    k = len(out_q)
    needs_perm = out_q != sorted(out_q)
    mapping = list(zip([k - q - 1 for q in out_q], [k - m - 1 for m in range(k)])) if needs_perm else []

    def flips(ix):
        p = _pad_bits(ix, endian)
        out = []
        for c in combos:
            f = p
            for b in c:
                f ^= 1 << b
            out.append(f)
        return out

    if not dm:
        for ix in range(D):
            for jx in range(ix, D):
                fi, fj = flips(ix), flips(jx)
                oi = _swap_bits(ix, mapping) if needs_perm else ix
                oj = _swap_bits(jx, mapping) if needs_perm else jx
                traced = complex(np.dot(np.conj(state[fj]), state[fi]))
                rho[oi, oj] += traced
                if oj != oi:
                    rho[oj, oi] += np.conj(traced)
    else:
        for raw in range(D * D):
            ix, jx = raw // D, raw % D
            fi, fj = flips(ix), flips(jx)
            oi = _swap_bits(ix, mapping) if needs_perm else ix
            oj = _swap_bits(jx, mapping) if needs_perm else jx
            rho[oi, oj] = sum(state[a, b] for a, b in zip(fi, fj))
    return rho


def permute_density_matrix(rho: np.ndarray, n: int, targets) -> np.ndarray:
    """src/result_types.jl:129-149."""
    t = list(targets)
    D = 2 ** len(t)
    out = np.zeros((D, D), dtype=rho.dtype)

    def perm(ix):
        bits = [(ix >> q) & 1 for q in range(n)]
        permed = [bits[n - tt - 1] for tt in t]
        return sum(b << q for q, b in enumerate(permed))

    for ix in range(D):
        for jx in range(D):
            out[perm(ix), perm(jx)] = rho[ix, jx]
    return out


# ---------------------------------------------------------------------------
# sampling checkers
# ---------------------------------------------------------------------------

def alias_reference(p: np.ndarray, wsum: float):
    """make_alias_table! (src/result_types.jl:8-54); alias returned 0-based."""
    n = p.size
    ap = np.empty(n)
    alias = np.zeros(n, dtype=np.int64)
    lw = np.zeros(n, dtype=np.int64)
    sw = np.zeros(n, dtype=np.int64)
    lib().ref_make_alias_table(_d(np.ascontiguousarray(p)), n, float(wsum), _d(ap), alias.ctypes.data_as(_i64p),
                               lw.ctypes.data_as(_i64p), sw.ctypes.data_as(_i64p))
    return ap, alias - 1


def direct_samples(p: np.ndarray, wsum: float, u: np.ndarray) -> np.ndarray:
    """StatsBase.sample per shot, the reference's n >= 30 sampler
    (src/result_types.jl:73-78): a linear scan of the running weight sum to
    u * wsum; 0-based shots."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty(u.size, dtype=np.int64)
    lib().ref_direct_sample(_d(np.ascontiguousarray(p)), p.size, float(wsum), u.size, _d(u), out.ctypes.data_as(_i64p))
    return out


def alias_exact(p: np.ndarray, wsum: float):
    n = p.size
    ap = np.empty(n)
    alias = np.empty(n, dtype=np.int32)
    lib().ref_alias_exact(_d(np.ascontiguousarray(p)), n, float(wsum), _d(ap), alias.ctypes.data_as(_i32p))
    return ap, alias


def shot_draws(seed: int, nq: int, n: int, first: int = 0):
    idx = np.empty(n, dtype=np.int64)
    u = np.empty(n)
    lib().ref_shot_draws(C.c_uint64(seed), nq, first, n, idx.ctypes.data_as(_i64p), _d(u))
    return idx, u


def cdf_shots(psi: np.ndarray, nq: int, shots: int, seed: int) -> np.ndarray:
    """Batched-execution shot rule: inverse CDF over exact integer weights."""
    out = np.empty(shots, dtype=np.int64)
    a = np.ascontiguousarray(psi, dtype=np.complex128)
    lib().ref_cdf_shots(_d(a), nq, shots, C.c_uint64(seed), out.ctypes.data_as(_i64p))
    return out


def lookup(ap: np.ndarray, alias: np.ndarray, idx: np.ndarray, u: np.ndarray) -> np.ndarray:
    """Shot rule of src/result_types.jl:71: u < ap[i] ? i : alias[i]."""
    return np.where(u < ap[idx], idx, alias[idx]).astype(np.int64)


def philox(ctr, key):
    i = (C.c_uint32 * 4)(*ctr)
    o = (C.c_uint32 * 4)()
    lib().ref_philox4x32_10(i, key[0], key[1], o)
    return list(o)


def splitmix_doubles(seed: int, n: int) -> np.ndarray:
    out = np.empty(n)
    lib().ref_splitmix64_doubles(C.c_uint64(seed), n, _d(out))
    return out


def alias_distribution(ap: np.ndarray, alias: np.ndarray) -> np.ndarray:
    """The distribution an alias table samples: (ap_k + sum over buckets b
    aliasing to k of (1 - ap_b)) / N."""
    n = ap.size
    q = ap.astype(np.float64).copy()
    np.add.at(q, alias, 1.0 - ap)
    return q / n


# ---------------------------------------------------------------------------
# numpy brute force (independent of refcpu)
# ---------------------------------------------------------------------------

def np_apply(psi: np.ndarray, n: int, m: np.ndarray, qubits) -> np.ndarray:
    """Apply ``m`` (kron order: qubits[0] is its most significant bit) to
    the qubits of a state whose qubit 0 is the most significant bit."""
    k = len(qubits)
    t = psi.reshape([2] * n)
    mt = np.asarray(m, dtype=np.complex128).reshape([2] * (2 * k))
    out = np.tensordot(mt, t, axes=(list(range(k, 2 * k)), list(qubits)))
    rest = [q for q in range(n) if q not in qubits]
    order = list(qubits) + rest
    inv = np.argsort(order)
    return np.transpose(out, inv).reshape(-1)


def np_controlled(m: np.ndarray, nc: int, cvals) -> np.ndarray:
    """Full matrix of a gate with nc leading controls (kron order)."""
    kt = int(round(math.log2(m.shape[0])))
    d = 2 ** (nc + kt)
    full = np.eye(d, dtype=np.complex128)
    idx = 0
    for b, v in enumerate(cvals):
        idx |= int(v) << (nc + kt - 1 - b)
    full[idx:idx + 2 ** kt, idx:idx + 2 ** kt] = m
    return full


def qft_state(thetas: np.ndarray) -> np.ndarray:
    """RY(theta_q) on every qubit of |0...0>, then the textbook QFT with
    qubit 0 as the most significant bit: sqrt(N) * ifft(a)."""
    a = np.ones(1, dtype=np.complex128)
    for th in thetas:
        a = np.kron(a, np.array([math.cos(th / 2), math.sin(th / 2)], dtype=np.complex128))
    return math.sqrt(a.size) * np.fft.ifft(a)
