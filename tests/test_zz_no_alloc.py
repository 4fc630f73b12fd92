"""bsim_evolve_noalloc (fused evolution with the plan tables in a fixed
device array instead of an allocated workspace) against bsim_evolve on the
same circuits, through the C ABI: two calls in a row, the second on a
smaller input, each bit-identical to the allocating entry point; plus the
refusals (ops outside the tile passes, tables beyond the 4 MiB scratch)."""
import ctypes
import sys

import numpy as np
import pytest

from bsim import _lib as L
from bsim import gates as G
from bsim import noises as Nz
from bsim import workloads as W
from bsim.simulators import DensityMatrixSimulator, Instruction, StateVectorSimulator, lower_instructions

from circuits import random_circuit, random_unitary

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(sim_cls, n, ins, **kw):
    """The same circuit through bsim_evolve (fused) and bsim_evolve_noalloc."""
    lib = L.load()
    a, b = sim_cls(n, 0, **kw), sim_cls(n, 0, **kw)
    ol = lower_instructions(ins, a.kind, kw.get("noise_mode", "reference"))
    L.call("bsim_evolve", a.handle, ol.ptr, ol.n, 0)
    rc = lib.bsim_evolve_noalloc(b.handle, ol.ptr, ol.n)
    assert rc == 0, L.last_error()
    return a._download(), b._download()


def test_noalloc_matches_evolve_twice_shrinking():
    FUS = ("1q", "2q", "ctrl", "diag")
    for n, depth, seed in ((20, 12, 1), (13, 5, 2)):  # second call: smaller input
        _say(f"no-alloc evolve: brick-wall n={n} depth={depth}")
        x, y = _pair(StateVectorSimulator, n, W.brickwall(n, depth, seed=seed))
        assert np.array_equal(x, y)
        rng = np.random.default_rng(seed)
        _say(f"no-alloc evolve: random fusable circuit n={n}")
        x, y = _pair(StateVectorSimulator, n, random_circuit(n, 120, rng, kinds=FUS))
        assert np.array_equal(x, y)


def test_noalloc_density_matrix_physical():
    rng = np.random.default_rng(5)
    n = 5
    ops = []
    for k in range(8):
        ops += random_circuit(n, 3, rng, kinds=("1q", "2q", "ctrl"))
        q = int(rng.integers(n))
        ops.append(Instruction([Nz.Depolarizing(0.05), Nz.AmplitudeDamping(0.1), Nz.PhaseFlip(0.2),
                                Nz.TwoQubitDepolarizing(0.03)][k % 4], [q] if k % 4 < 3 else [q, (q + 2) % n]))
    _say("no-alloc evolve: 5-qubit noisy density matrix, physical channels")
    x, y = _pair(DensityMatrixSimulator, n, ops, noise_mode="physical")
    assert np.array_equal(x, y)


def test_noalloc_refusals():
    lib = L.load()
    sim = StateVectorSimulator(8, 0)
    u3 = random_unitary(3, np.random.default_rng(0))
    ol = lower_instructions([Instruction(G.H(), [0]), Instruction(G.Unitary(u3), [0, 1, 2])], L.BSIM_SV)
    assert lib.bsim_evolve_noalloc(sim.handle, ol.ptr, ol.n) == L.BSIM_ENOTSUP
    assert "fusable" in L.last_error()
    # 72,000 gates: tens of thousands of tile ops do not fit the 4 MiB tables
    big = lower_instructions(W.brickwall(12, 4000, seed=3), L.BSIM_SV)
    sim12 = StateVectorSimulator(12, 0)
    assert lib.bsim_evolve_noalloc(sim12.handle, big.ptr, big.n) == L.BSIM_ENOTSUP
    assert "4 MiB" in L.last_error()
    # the handle still evolves after the refusals
    small = lower_instructions([Instruction(G.X(), [0])], L.BSIM_SV)
    assert lib.bsim_evolve_noalloc(sim.handle, small.ptr, small.n) == 0
    p = sim.probabilities()
    assert p[128] == 1.0
