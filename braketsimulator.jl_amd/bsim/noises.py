"""Noise channels and their Kraus representations (src/noises.jl,
src/noise_kernels.jl:25-103)."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .gates import I2, PX, PY, PZ


class Noise:
    n_qubits = 1
    name = "Noise"

    def kraus_rep(self) -> list:
        raise NotImplementedError


@dataclass
class Kraus(Noise):
    matrices: list
    name = "Kraus"

    def __post_init__(self):
        self.matrices = [np.asarray(m, dtype=np.complex128) for m in self.matrices]
        self.n_qubits = int(round(math.log2(self.matrices[0].shape[0])))

    def kraus_rep(self):
        return list(self.matrices)


@dataclass
class BitFlip(Noise):
    probability: float
    name = "BitFlip"

    def kraus_rep(self):  # :36
        p = self.probability
        return [math.sqrt(1 - p) * I2, math.sqrt(p) * PX]


@dataclass
class PhaseFlip(Noise):
    """Applied by its own elementwise kernel (src/noise_kernels.jl:7-23)."""

    probability: float
    name = "PhaseFlip"

    def kraus_rep(self):
        p = self.probability
        return [math.sqrt(1 - p) * I2, math.sqrt(p) * PZ]


@dataclass
class Depolarizing(Noise):
    probability: float
    name = "Depolarizing"

    def kraus_rep(self):  # :30-35
        p = self.probability
        return [math.sqrt(1.0 - p) * I2] + [math.sqrt(p / 3.0) * m for m in (PX, PY, PZ)]


@dataclass
class PauliChannel(Noise):
    probX: float
    probY: float
    probZ: float
    name = "PauliChannel"

    def kraus_rep(self):  # :37-42
        return [
            math.sqrt(1 - self.probX - self.probY - self.probZ) * I2,
            math.sqrt(self.probX) * PX,
            math.sqrt(self.probY) * PY,
            math.sqrt(self.probZ) * PZ,
        ]


@dataclass
class AmplitudeDamping(Noise):
    gamma: float
    name = "AmplitudeDamping"

    def kraus_rep(self):  # :43
        g = self.gamma
        return [np.array([[1.0, 0.0], [0.0, math.sqrt(1.0 - g)]], dtype=np.complex128),
                np.array([[0.0, math.sqrt(g)], [0.0, 0.0]], dtype=np.complex128)]


@dataclass
class GeneralizedAmplitudeDamping(Noise):
    gamma: float
    probability: float
    name = "GeneralizedAmplitudeDamping"

    def kraus_rep(self):  # :44-49
        p, g = self.probability, self.gamma
        return [
            math.sqrt(p) * np.array([[1.0, 0.0], [0.0, math.sqrt(1.0 - g)]], dtype=np.complex128),
            math.sqrt(p) * np.array([[0.0, math.sqrt(g)], [0.0, 0.0]], dtype=np.complex128),
            math.sqrt(1.0 - p) * np.array([[math.sqrt(1.0 - g), 0.0], [0.0, 1.0]], dtype=np.complex128),
            math.sqrt(1.0 - p) * np.array([[0.0, 0.0], [math.sqrt(g), 0.0]], dtype=np.complex128),
        ]


@dataclass
class PhaseDamping(Noise):
    gamma: float
    name = "PhaseDamping"

    def kraus_rep(self):  # :50
        g = self.gamma
        return [np.array([[1.0, 0.0], [0.0, math.sqrt(1.0 - g)]], dtype=np.complex128),
                np.array([[0.0, 0.0], [0.0, math.sqrt(g)]], dtype=np.complex128)]


@dataclass
class TwoQubitDepolarizing(Noise):
    probability: float
    name = "TwoQubitDepolarizing"
    n_qubits = 2

    def kraus_rep(self):  # :52-64 (column-major vec of [kron(ki, kj) for ki, kj])
        p = self.probability
        ks = (I2, PX, PY, PZ)
        out = []
        for kj in ks:          # Julia comprehension [f(ki,kj) for ki in ks, kj in ks] is column-major
            for ki in ks:
                out.append(np.kron(ki, kj))
        out[0] = out[0] * math.sqrt(1.0 - p)
        f = math.sqrt(p / 15.0)
        return [out[0]] + [m * f for m in out[1:]]


@dataclass
class TwoQubitDephasing(Noise):
    probability: float
    name = "TwoQubitDephasing"
    n_qubits = 2

    def kraus_rep(self):  # :80-100
        p = self.probability
        return [
            math.sqrt(1 - p) * np.kron(I2, I2),
            math.sqrt(p / 3.0) * np.kron(I2, PZ),
            math.sqrt(p / 3.0) * np.kron(PZ, I2),
            math.sqrt(p / 3.0) * np.kron(PZ, PZ),
        ]


@dataclass
class TwoQubitPauliChannel(Noise):
    probabilities: dict
    name = "TwoQubitPauliChannel"
    n_qubits = 2

    def kraus_rep(self):  # :66-78
        d = {"I": I2, "X": PX, "Y": PY, "Z": PZ}
        total = sum(self.probabilities.values())
        ks = [math.sqrt(1.0 - total) * np.eye(4, dtype=np.complex128)]
        for k, v in self.probabilities.items():
            if k != "II":
                ks.append(math.sqrt(v) * np.kron(d[k[0]], d[k[1]]))
        return ks


class MultiQubitPauliChannel(TwoQubitPauliChannel):
    """MultiQubitPauliChannel{N} (src/noises.jl:71-90): N is the length of
    the Pauli words in ``probabilities``.  The reference defines Kraus
    matrices for N = 2 only (TwoQubitPauliChannel, src/noise_kernels.jl:66);
    other N raise, as there."""

    name = "MultiQubitPauliChannel"

    def __init__(self, probabilities: dict):
        self.probabilities = dict(probabilities)
        self.n_qubits = len(next(iter(self.probabilities))) if self.probabilities else 0

    def kraus_rep(self):
        if self.n_qubits != 2:
            raise NotImplementedError("Kraus matrices of a Pauli channel are defined for 2 qubits only")
        return TwoQubitPauliChannel.kraus_rep(self)


NOISES = {c.name: c for c in (Kraus, BitFlip, PhaseFlip, Depolarizing, PauliChannel, AmplitudeDamping,
                              GeneralizedAmplitudeDamping, PhaseDamping, TwoQubitDepolarizing,
                              TwoQubitDephasing, TwoQubitPauliChannel, MultiQubitPauliChannel)}
