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
    """n doubles in [0, 1) from SplitMix64 seeded with ``seed``."""
    out = np.empty(n, dtype=np.float64)
    x = seed & _M64
    for i in range(n):
        x = (x + 0x9E3779B97F4A7C15) & _M64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        z ^= z >> 31
        out[i] = (z >> 11) * 2.0 ** -53
    return out


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
