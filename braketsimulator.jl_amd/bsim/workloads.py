"""Benchmark circuits: the RY+QFT circuit, the brick-wall random circuit
(state vector and noisy density matrix), the hardware-efficient ansatz with
its transverse-field Ising Hamiltonian, and the uneven batch.  Angles come
from SplitMix64 (Steele, Lea, Flood 2014), double = (x >> 11) * 2^-53."""
from __future__ import annotations

import math

import numpy as np

from . import gates as G
from . import noises as Nz
from .simulators import Instruction

_M64 = (1 << 64) - 1


def splitmix64_doubles(seed: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from SplitMix64 seeded with ``seed`` (state i is
    seed + (i + 1) * 0x9E3779B97F4A7C15 mod 2^64, mixed)."""
    with np.errstate(over="ignore"):
        x = np.uint64(seed & _M64) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def ry_qft(n: int = 20, seed: int = 0):
    """RY(theta_q), theta ~ U[0, pi), on every qubit, then the textbook QFT:
    H(j), CPhaseShift(pi / 2^(k-j)) on [k, j] for k > j, then swaps
    j <-> n-1-j.  Returns (instructions, thetas)."""
    th = splitmix64_doubles(seed, n) * math.pi
    ins = [Instruction(G.Ry(float(t)), [q]) for q, t in enumerate(th)]
    for j in range(n):
        ins.append(Instruction(G.H(), [j]))
        for k in range(j + 1, n):
            ins.append(Instruction(G.CPhaseShift(math.pi / 2 ** (k - j)), [k, j]))
    for j in range(n // 2):
        ins.append(Instruction(G.Swap(), [j, n - 1 - j]))
    return ins, th


def brickwall(n: int = 30, depth: int = 20, seed: int = 0, noise: bool = False):
    """Layer l: U(theta, phi, lambda) ~ U[0, 2 pi)^3 on every qubit, then CZ on
    (0,1),(2,3),... (even l) or (1,2),(3,4),... (odd l).  With ``noise``:
    Depolarizing(0.01) after every U, TwoQubitDepolarizing(0.02) after every
    CZ and AmplitudeDamping(0.005) on every qubit at the end of each layer."""
    ang = splitmix64_doubles(seed, 3 * n * depth) * (2 * math.pi)
    ins = []
    a = 0
    for layer in range(depth):
        for q in range(n):
            ins.append(Instruction(G.U(float(ang[a]), float(ang[a + 1]), float(ang[a + 2])), [q]))
            a += 3
            if noise:
                ins.append(Instruction(Nz.Depolarizing(0.01), [q]))
        for q in range(layer % 2, n - 1, 2):
            ins.append(Instruction(G.CZ(), [q, q + 1]))
            if noise:
                ins.append(Instruction(Nz.TwoQubitDepolarizing(0.02), [q, q + 1]))
        if noise:
            for q in range(n):
                ins.append(Instruction(Nz.AmplitudeDamping(0.005), [q]))
    return ins


def hea(n: int = 24, layers: int = 8, seed: int = 0, noise_p: float = 0.0):
    """Hardware-efficient ansatz: per layer RY(theta) RZ(theta) on every qubit,
    then CNOT 0->1->...->n-1.  theta ~ U[-pi, pi).  Returns (instructions,
    param_map, thetas): param_map[(instruction index, 0)] = parameter index."""
    th = splitmix64_doubles(seed, 2 * n * layers) * (2 * math.pi) - math.pi
    ins, pmap = [], {}
    p = 0
    for _ in range(layers):
        for q in range(n):
            for gate in (G.Ry, G.Rz):
                pmap[(len(ins), 0)] = p
                ins.append(Instruction(gate(float(th[p])), [q]))
                p += 1
                if noise_p:
                    ins.append(Instruction(Nz.Depolarizing(noise_p), [q]))
        for q in range(n - 1):
            ins.append(Instruction(G.CNot(), [q, q + 1]))
            if noise_p:
                ins.append(Instruction(Nz.TwoQubitDepolarizing(noise_p), [q, q + 1]))
    return ins, pmap, th


def tfim(n: int = 24, hx: float = 0.5):
    """Open-chain transverse-field Ising H = sum Z_i Z_{i+1} + hx sum X_i as
    (Pauli words, coefficients)."""
    words, coeffs = [], []
    for i in range(n - 1):
        w = ["I"] * n
        w[i] = w[i + 1] = "Z"
        words.append("".join(w))
        coeffs.append(1.0)
    for i in range(n):
        w = ["I"] * n
        w[i] = "X"
        words.append("".join(w))
        coeffs.append(hx)
    return words, np.array(coeffs)


def batch_tasks(count: int = 4096, seed: int = 0, qmin: int = 8, qmax: int = 22, dmin: int = 10, dmax: int = 40):
    """(qubit count, depth, circuit seed) for the uneven batch."""
    u = splitmix64_doubles(seed ^ 0x5EED, 2 * count)
    out = []
    for i in range(count):
        n = qmin + int(u[2 * i] * (qmax - qmin + 1))
        d = dmin + int(u[2 * i + 1] * (dmax - dmin + 1))
        out.append((n, d, seed + 1 + i))
    return out


class PackedOps:
    """A bsim_op array built directly in numpy (same layout as the ctypes
    ``Op``), with the matrix buffer it points into; ``n`` and ``ptr`` as for
    ``OpList``."""

    def __init__(self, rec: np.ndarray, mats: np.ndarray):
        self.rec, self.mats = rec, mats
        self.n = len(rec)
        self.addr = rec.__array_interface__["data"][0]  # address of the first record

    @property
    def ptr(self):
        import ctypes as C

        from . import _lib as L
        return C.cast(C.c_void_p(self.rec.ctypes.data), C.POINTER(L.Op))


def _op_dtype():
    from . import _lib as L
    f = {name: getattr(L.Op, name).offset for name, _ in L.Op._fields_}
    return np.dtype({"names": ["type", "n_targets", "n_controls", "target0", "control0", "cval0", "matrix"],
                     "formats": ["<i4", "<i4", "<i4", "<i4", "<i4", "<i4", "<u8"],
                     "offsets": [f["type"], f["n_targets"], f["n_controls"], f["targets"], f["controls"],
                                 f["control_values"], f["matrix"]],
                     "itemsize": __import__("ctypes").sizeof(L.Op)})


def brickwall_ops(n: int, depth: int, seed: int = 0) -> PackedOps:
    """``brickwall(n, depth, seed)`` lowered straight into a bsim_op array:
    U(theta, phi, lambda) = [[c, -e^{i lam} s], [e^{i phi} s, e^{i(phi+lam)} c]]
    (c, s = cos, sin of theta/2) as a 1-target gate, CZ(q, q+1) as Z on q+1
    controlled by q -- the kernel calls ``lower_instructions`` makes."""
    from . import _lib as L
    ang = splitmix64_doubles(seed, 3 * n * depth).reshape(depth, n, 3) * (2 * math.pi)
    th, ph, lam = ang[..., 0], ang[..., 1], ang[..., 2]
    c, s = np.cos(th / 2), np.sin(th / 2)
    m = np.empty((depth, n, 4), dtype=np.complex128)   # column-major 2x2
    m[..., 0] = c
    m[..., 1] = np.exp(1j * ph) * s
    m[..., 2] = -np.exp(1j * lam) * s
    m[..., 3] = np.exp(1j * (lam + ph)) * c
    mats = np.concatenate([m.reshape(-1, 4), np.array([[1, 0, 0, -1]], dtype=np.complex128)])
    zaddr = mats.ctypes.data + 64 * (len(mats) - 1)
    count = depth * n + sum(len(range(l % 2, n - 1, 2)) for l in range(depth))
    rec = np.zeros(count, dtype=_op_dtype())
    k = 0
    for layer in range(depth):
        sl = slice(k, k + n)
        rec["type"][sl] = L.OP_GATE
        rec["n_targets"][sl] = 1
        rec["target0"][sl] = np.arange(n)
        rec["matrix"][sl] = mats.ctypes.data + 64 * (layer * n + np.arange(n, dtype=np.uint64))
        k += n
        qs = np.arange(layer % 2, n - 1, 2)
        sl = slice(k, k + len(qs))
        rec["type"][sl] = L.OP_GATE
        rec["n_targets"][sl] = 1
        rec["n_controls"][sl] = 1
        rec["target0"][sl] = qs + 1
        rec["control0"][sl] = qs
        rec["cval0"][sl] = 1
        rec["matrix"][sl] = zaddr
        k += len(qs)
    return PackedOps(rec, mats)
