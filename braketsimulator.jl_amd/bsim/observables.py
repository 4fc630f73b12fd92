"""Observables (src/observables.jl): the standard Paulis, H, I, tensor
products and Hermitian matrices."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .gates import I2, PX, PY, PZ, SQ2


class Observable:
    n_qubits = 1
    pauli = None  # letter for Pauli observables

    def matrix(self) -> np.ndarray:
        raise NotImplementedError


class X(Observable):
    pauli = "X"

    def matrix(self):
        return PX.copy()


class Y(Observable):
    pauli = "Y"

    def matrix(self):
        return PY.copy()


class Z(Observable):
    pauli = "Z"

    def matrix(self):
        return PZ.copy()


class I(Observable):
    pauli = "I"

    def matrix(self):
        return I2.copy()


class H(Observable):
    def matrix(self):
        return SQ2 * np.array([[1, 1], [1, -1]], dtype=np.complex128)


@dataclass
class HermitianObservable(Observable):
    mat: np.ndarray

    def __post_init__(self):
        self.mat = np.asarray(self.mat, dtype=np.complex128)
        self.n_qubits = int(round(math.log2(self.mat.shape[0])))

    def matrix(self):
        return self.mat.copy()


@dataclass
class TensorProduct(Observable):
    factors: list

    def __post_init__(self):
        self.n_qubits = sum(f.n_qubits for f in self.factors)

    def matrix(self):
        m = np.ones((1, 1), dtype=np.complex128)
        for f in self.factors:
            m = np.kron(m, f.matrix())
        return m


def is_pauli_word(obs: Observable) -> bool:
    if obs.pauli is not None:
        return True
    return isinstance(obs, TensorProduct) and all(f.pauli is not None for f in obs.factors)


def pauli_string(obs: Observable, targets, n_qubits: int) -> str:
    """Full-register Pauli word ('I','X','Y','Z' per qubit) of a Pauli
    observable on ``targets``."""
    word = ["I"] * n_qubits
    if obs.pauli is not None:
        for t in targets:
            word[t] = obs.pauli
        return "".join(word)
    for f, t in zip(obs.factors, targets):
        word[t] = f.pauli
    return "".join(word)
