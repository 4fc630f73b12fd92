"""Result types and ``calculate`` (src/results.jl, src/result_types.jl:82-246)."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .observables import Observable


@dataclass
class Expectation:
    observable: Observable
    targets: list = field(default_factory=list)


@dataclass
class Variance:
    observable: Observable
    targets: list = field(default_factory=list)


@dataclass
class Probability:
    targets: list = field(default_factory=list)


@dataclass
class Sample:
    observable: Observable
    targets: list = field(default_factory=list)


@dataclass
class StateVector:
    pass


@dataclass
class DensityMatrix:
    targets: list = field(default_factory=list)


@dataclass
class Amplitude:
    states: list


def permute_probability(probs: np.ndarray, n: int, targets) -> np.ndarray:
    """src/result_types.jl:115-127."""
    out = np.zeros_like(probs)
    t = list(targets)
    for ix in range(2 ** len(t)):
        unperm = [(ix >> q) & 1 for q in range(n)]
        perm = [unperm[n - tt - 1] for tt in t]  # unpermed_bits[qubit_count .- targets] (1-based)
        new = 0
        for b, q in zip(perm, range(n)):
            new += b << q
        out[new] = probs[ix]
    return out


def permute_density_matrix(rho: np.ndarray, n: int, targets) -> np.ndarray:
    """src/result_types.jl:129-149: new bit q of an index = the bit of
    qubit targets[q]."""
    t = list(targets)
    D = 2 ** len(t)

    def perm(ix):
        unperm = [(ix >> q) & 1 for q in range(n)]
        return sum(unperm[n - tt - 1] << q for q, tt in enumerate(t))

    p = [perm(ix) for ix in range(D)]
    out = np.zeros((D, D), dtype=rho.dtype)
    for ix in range(D):
        out[p[ix], p] = rho[ix, :]
    return out


def calculate(rt, sim):
    """calculate(result_type, simulator)."""
    n = sim.qubit_count
    if isinstance(rt, Probability):
        t = list(rt.targets or [])
        if not t or t == list(range(n)):
            return sim.probabilities()
        if len(t) == n:
            return permute_probability(sim.probabilities(), n, t)
        return sim.marginal_probabilities(t)
    if isinstance(rt, Expectation):
        t = list(rt.targets) if rt.targets else list(range(n))
        if len(t) == rt.observable.n_qubits:
            return sim.expectation(rt.observable, *t)
        if rt.observable.pauli == "Z" and t == list(range(n)):
            return list(sim.expect_z_all())  # one pass for every <Z_q>
        return [sim.expectation(rt.observable, q) for q in t]
    if isinstance(rt, Variance):
        t = list(rt.targets) if rt.targets else list(range(n))
        if len(t) == rt.observable.n_qubits:
            return sim.expectation_op_squared(rt.observable, *t) - sim.expectation(rt.observable, *t) ** 2
        return [sim.expectation_op_squared(rt.observable, q) - sim.expectation(rt.observable, q) ** 2 for q in t]
    if isinstance(rt, StateVector):
        return sim.state_vector()
    if isinstance(rt, DensityMatrix):  # src/result_types.jl:240-246
        t = list(rt.targets or [])
        if not t or t == list(range(n)):
            return sim.density_matrix()
        if len(t) == n:
            return permute_density_matrix(sim.density_matrix(), n, t)
        return sim.partial_trace(t)
    if isinstance(rt, Amplitude):
        sv = sim.state_vector()
        return {s: sv[int(s, 2)] for s in rt.states}
    raise TypeError(f"unsupported result type {rt!r}")
