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


_TASK_DTYPE = None


def _task_dtype():
    """numpy layout of bsim_task (the ctypes ``Task``): the batch's task
    array is filled by vector operations instead of one ctypes assignment
    per field and task."""
    global _TASK_DTYPE
    if _TASK_DTYPE is None:
        import ctypes as C

        names = [n for n, _ in L.Task._fields_]
        fmt = {"qubit_count": "<i4", "n_ops": "<i4", "shots": "<i8", "n_words": "<i4", "n_params": "<i4"}
        _TASK_DTYPE = np.dtype({"names": names, "formats": [fmt.get(n, "<u8") for n in names],
                                "offsets": [getattr(L.Task, n).offset for n in names],
                                "itemsize": C.sizeof(L.Task)})
    return _TASK_DTYPE


def _is_op_array(x) -> bool:
    """An already lowered op array (OpList, PackedOps): tested on the type,
    since reading ``ptr`` on an OpList builds its record array."""
    return hasattr(type(x), "ptr")


def _address(ol) -> int:
    """Address of an op array (OpList, PackedOps or anything with ``ptr``)."""
    import ctypes as C

    a = getattr(ol, "addr", None)
    if isinstance(a, int):
        return a
    return C.cast(ol.ptr, C.c_void_p).value or 0


def run_batch(circuits, shots=1000, seeds=None):
    """circuits: list of (qubit_count, instructions) -- or (qubit_count, ops)
    with ops already a bsim_op array (``n`` and ``ptr``, as OpList and
    workloads.PackedOps have) -- and GradientTask items, all run in the same
    persistent launch.  ``shots`` and ``seeds`` are scalars or per-item
    sequences (gradient tasks take no shots).  Returns a list of BatchResult."""
    import ctypes as C

    n = len(circuits)
    if n == 0:
        return []
    shots_a = np.broadcast_to(np.asarray(shots, dtype=np.int64), (n,)).copy()
    if seeds is None:
        seeds_a = np.arange(n, dtype=np.uint64)
    elif np.ndim(seeds):
        seeds_a = np.array([int(s) & 0xFFFFFFFFFFFFFFFF for s in seeds], dtype=np.uint64)
    else:
        seeds_a = (np.uint64(int(seeds) & 0xFFFFFFFFFFFFFFFF) + np.arange(n, dtype=np.uint64))
    tasks = np.zeros(n, dtype=_task_dtype())
    tasks["seed"] = seeds_a
    is_grad = np.array([isinstance(c, GradientTask) for c in circuits], dtype=bool)
    keep = []
    nq = np.zeros(n, dtype=np.int64)
    nops = np.zeros(n, dtype=np.int64)
    addr = np.zeros(n, dtype=np.uint64)
    grads = {}
    for i in np.flatnonzero(is_grad):
        item = circuits[i]
        q = item.qubit_count
        pm = item.param_map
        npar = (max(pm.values()) + 1 if pm else 0) if item.n_params is None else int(item.n_params)
        ol = item.instructions if _is_op_array(item.instructions) else \
            lower_instructions(list(item.instructions), L.BSIM_SV, param_map=pm)
        for w in item.words:
            if len(w) != q:
                raise ValueError(f"Pauli word {w!r} must have {q} letters")
        buf = C.create_string_buffer("".join(item.words).encode())
        c = np.ascontiguousarray(item.coeffs, dtype=np.float64)
        g = np.zeros(max(npar, 1), dtype=np.float64)
        ev = np.zeros(1, dtype=np.float64)
        keep.append((ol, buf, c))
        grads[i] = (g[:npar], ev)
        nq[i], nops[i], addr[i] = q, ol.n, _address(ol)
        t = tasks[i]
        t["shots"] = 0
        t["n_words"] = len(item.words)
        t["n_params"] = npar
        t["paulis"] = C.addressof(buf)
        t["coeffs"] = c.ctypes.data
        t["grad"] = g.ctypes.data
        t["expval"] = ev.ctypes.data
    circ = np.flatnonzero(~is_grad)
    for i in circ:
        q, ins = circuits[i]
        ol = ins if _is_op_array(ins) else lower_instructions(ins, L.BSIM_SV)
        keep.append(ol)
        nq[i], nops[i], addr[i] = q, ol.n, _address(ol)
    # the circuits' outputs in two blocks: <Z_q> of every circuit, then shots
    shots_a[is_grad] = 0
    zoff = np.zeros(n + 1, dtype=np.int64)
    zoff[1:] = np.cumsum(np.where(is_grad, 0, nq))
    soff = np.zeros(n + 1, dtype=np.int64)
    soff[1:] = np.cumsum(shots_a)
    z = np.zeros(max(int(zoff[-1]), 1), dtype=np.float64)
    s = np.zeros(max(int(soff[-1]), 1), dtype=np.int64)
    tasks["qubit_count"] = nq
    tasks["n_ops"] = nops
    tasks["ops"] = addr
    tasks["shots"] = shots_a
    tasks["expect_z"][circ] = np.uint64(z.ctypes.data) + (zoff[:-1][circ] * 8).astype(np.uint64)
    tasks["samples"][circ] = np.where(shots_a[circ] > 0,
                                      np.uint64(s.ctypes.data) + (soff[:-1][circ] * 8).astype(np.uint64),
                                      np.uint64(0))
    lat = np.zeros(n, dtype=np.float64)
    L.call("bsim_batch_run", tasks.ctypes.data_as(C.POINTER(L.Task)), n, L.dptr(lat))
    del keep
    res = []
    empty_z, empty_s = np.zeros(0), np.zeros(0, dtype=np.int64)
    for i in range(n):
        if is_grad[i]:
            g, ev = grads[i]
            res.append(BatchResult(empty_z, empty_s, float(lat[i]), g.copy(), float(ev[0])))
        else:
            res.append(BatchResult(z[zoff[i]:zoff[i + 1]], s[soff[i]:soff[i] + shots_a[i]], float(lat[i])))
    return res
