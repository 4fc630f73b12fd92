"""Batched state-vector execution in one persistent launch.

The reference runs a batch as one ``simulate`` per circuit on a pool of
``min(32, nthreads)`` Julia tasks (src/BraketSimulator.jl:347-395).  Here the
whole batch is one call of ``bsim_batch_run``: every circuit is planned into
fused tile passes and a persistent kernel pulls (circuit, pass, tile)
tickets from a device work queue, so circuits of uneven qubit count and depth
share the GPU.  Each circuit returns <Z_q> for every qubit and ``shots``
basis-state indices.

Shots here are drawn by inverse CDF over the exact integer weights
W_i = floor(p_i 2^n 2^32) with Philox-4x32-10(counter = shot index, key =
seed), r = (x1:x0), t = floor(r * sum(W) / 2^64): the first i with
W_0 + ... + W_i > t.  The distribution is the alias method's (both sample
the same probability vector), the individual draws are not.

Parameter sweeps ride in the same launch: a ``GradientTask`` runs the
circuit forward, forms lambda = H psi and sweeps backward with the fused
adjoint tile body, returning the gradient of bsim_adjoint_gradient.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .simulators import lower_instructions


@dataclass
class BatchResult:
    expect_z: np.ndarray      # <Z_q>, qubit q = index q (circuits)
    samples: np.ndarray       # basis-state indices (qubit 0 = MSB) (circuits)
    latency_us: float         # completion time from the launch start
    grad: np.ndarray | None = None   # gradient tasks: d<H>/dtheta
    expval: float | None = None      # gradient tasks: <H>


@dataclass
class GradientTask:
    """One parameter set of a sweep: the adjoint gradient of <H> =
    sum_w coeffs[w] <P_w> for ``instructions`` from |0...0>, with
    ``param_map`` as for adjoint.adjoint_gradient ((instruction index,
    angle index) -> parameter index).  ``instructions`` may be an op list
    already lowered with that map."""

    qubit_count: int
    instructions: list
    param_map: dict
    words: list
    coeffs: np.ndarray
    n_params: int | None = None


def run_batch(circuits, shots=1000, seeds=None):
    """circuits: list of (qubit_count, instructions) -- or (qubit_count, ops)
    with ops already a bsim_op array (``n`` and ``ptr``, as OpList and
    workloads.PackedOps have) -- and GradientTask items, all run in the same
    persistent launch.  ``shots`` and ``seeds`` are scalars or per-item
    sequences (gradient tasks take no shots).  Returns a list of BatchResult."""
    n = len(circuits)
    if n == 0:
        return []
    shots_l = list(shots) if np.ndim(shots) else [int(shots)] * n
    seeds_l = list(seeds) if seeds is not None and np.ndim(seeds) else [
        (0 if seeds is None else int(seeds)) + i for i in range(n)]
    tasks = (L.Task * n)()
    keep, outs = [], []
    for i, item in enumerate(circuits):
        t = tasks[i]
        t.seed = int(seeds_l[i]) & 0xFFFFFFFFFFFFFFFF
        if isinstance(item, GradientTask):
            nq = item.qubit_count
            pm = item.param_map
            npar = (max(pm.values()) + 1 if pm else 0) if item.n_params is None else int(item.n_params)
            if hasattr(item.instructions, "ptr"):  # lowered with param_map already
                ol = item.instructions
            else:
                ol = lower_instructions(list(item.instructions), L.BSIM_SV, param_map=pm)
            for w in item.words:
                if len(w) != nq:
                    raise ValueError(f"Pauli word {w!r} must have {nq} letters")
            buf = "".join(item.words).encode()
            c = np.ascontiguousarray(item.coeffs, dtype=np.float64)
            g = np.zeros(max(npar, 1), dtype=np.float64)
            ev = np.zeros(1, dtype=np.float64)
            keep.append((ol, buf, c))
            outs.append(("grad", g[:npar], ev))
            t.qubit_count = int(nq)
            t.n_ops = ol.n
            t.ops = ol.ptr
            t.shots = 0
            t.n_words = len(item.words)
            t.n_params = npar
            t.paulis = buf
            t.coeffs = L.dptr(c)
            t.grad = L.dptr(g)
            t.expval = L.dptr(ev)
            continue
        nq, ins = item
        ol = ins if hasattr(ins, "ptr") else lower_instructions(ins, L.BSIM_SV)
        keep.append(ol)
        z = np.zeros(nq, dtype=np.float64)
        s = np.zeros(max(int(shots_l[i]), 1), dtype=np.int64)
        outs.append(("circuit", z, s[: int(shots_l[i])]))
        t.qubit_count = int(nq)
        t.n_ops = ol.n
        t.ops = ol.ptr
        t.shots = int(shots_l[i])
        t.expect_z = L.dptr(z)
        t.samples = L.i64ptr(s)
    lat = np.zeros(n, dtype=np.float64)
    L.call("bsim_batch_run", tasks, n, L.dptr(lat))
    res = []
    for i, o in enumerate(outs):
        if o[0] == "grad":
            res.append(BatchResult(np.zeros(0), np.zeros(0, dtype=np.int64), float(lat[i]), o[1].copy(),
                                   float(o[2][0])))
        else:
            res.append(BatchResult(o[1], o[2], float(lat[i])))
    return res
