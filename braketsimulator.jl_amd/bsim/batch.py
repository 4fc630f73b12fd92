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
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .simulators import lower_instructions


@dataclass
class BatchResult:
    expect_z: np.ndarray      # <Z_q>, qubit q = index q
    samples: np.ndarray       # basis-state indices (qubit 0 = MSB)
    latency_us: float         # completion time from the launch start


def run_batch(circuits, shots=1000, seeds=None):
    """circuits: list of (qubit_count, instructions) -- or (qubit_count, ops)
    with ops already a bsim_op array (``n`` and ``ptr``, as OpList and
    workloads.PackedOps have).  ``shots`` and ``seeds`` are scalars or
    per-circuit sequences.  Returns a list of BatchResult."""
    n = len(circuits)
    if n == 0:
        return []
    shots_l = list(shots) if np.ndim(shots) else [int(shots)] * n
    seeds_l = list(seeds) if seeds is not None and np.ndim(seeds) else [
        (0 if seeds is None else int(seeds)) + i for i in range(n)]
    tasks = (L.Task * n)()
    keep, outs = [], []
    for i, (nq, ins) in enumerate(circuits):
        ol = ins if hasattr(ins, "ptr") else lower_instructions(ins, L.BSIM_SV)
        keep.append(ol)
        z = np.zeros(nq, dtype=np.float64)
        s = np.zeros(max(int(shots_l[i]), 1), dtype=np.int64)
        outs.append((z, s))
        t = tasks[i]
        t.qubit_count = int(nq)
        t.n_ops = ol.n
        t.ops = ol.ptr
        t.shots = int(shots_l[i])
        t.seed = int(seeds_l[i]) & 0xFFFFFFFFFFFFFFFF
        t.expect_z = L.dptr(z)
        t.samples = L.i64ptr(s)
    lat = np.zeros(n, dtype=np.float64)
    L.call("bsim_batch_run", tasks, n, L.dptr(lat))
    return [BatchResult(z, s[: int(shots_l[i])], float(lat[i])) for i, (z, s) in enumerate(outs)]
