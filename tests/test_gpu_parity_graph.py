"""Captured evolutions (bsim_graph_create / launch / destroy) and the tile
autotuner (bsim_autotune).

A replayed graph runs the same fused plan as bsim_evolve, so every replay
must be bit-identical to an eager evolve of the same circuit (and within
1e-12 of the oracle).  Replays are separated by eager work on the same and
on other handles, with synchronizations, and each one is checked.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle as O
from bsim import BsimError
from bsim import _lib as L
from bsim import gates as G
from bsim import noises as Nz
from bsim import workloads as W
from bsim.simulators import DensityMatrixSimulator, Instruction, StateVectorSimulator, lower_instructions

from circuits import random_circuit, random_unitary

pytestmark = pytest.mark.gpu

FUSABLE = ("1q", "2q", "ctrl", "diag")  # every gate lands in a tile pass (dense <= 2 targets)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture
def keep_tile_config():
    t, lo, rb, fl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L.call("bsim_get_tile_config", ctypes.byref(t), ctypes.byref(lo))
    L.call("bsim_set_sweep_bits", 0, ctypes.byref(rb))
    L.call("bsim_set_fusion_limit", 0, ctypes.byref(fl))
    L.call("bsim_set_fusion_limit", fl.value, None)
    mw = ctypes.c_int()
    L.call("bsim_set_merge_width", -1, ctypes.byref(mw))
    yield
    L.call("bsim_set_tile_config", t.value, lo.value)
    L.call("bsim_set_sweep_bits", rb.value, None)
    L.call("bsim_set_fusion_limit", fl.value, None)
    L.call("bsim_set_merge_width", mw.value, None)


def _plan_singles(n, kind, ins, mode="physical"):
    ol = lower_instructions(ins, kind, mode)
    st = (ctypes.c_int64 * 4)()
    vq = n if kind == L.BSIM_SV else 2 * n
    L.call("bsim_plan_stats", vq, int(kind == L.BSIM_DM), ol.ptr, ol.n, st)
    return st[1]


@pytest.mark.parametrize("n,depth", [(6, 4), (14, 10), (20, 6)])
def test_graph_replays_match_eager_sv(n, depth):
    ins = W.brickwall(n, depth, seed=n)
    assert _plan_singles(n, L.BSIM_SV, ins) == 0
    _say(f"graph sv brickwall n={n} depth={depth}")
    eager = StateVectorSimulator(n, 0).evolve(ins).state_vector()
    np.testing.assert_allclose(eager, O.evolve_sv(n, ins), atol=1e-12)
    sim = StateVectorSimulator(n, 0)
    other = StateVectorSimulator(n, 0)
    with sim.capture(ins, reset=True) as g:
        for rep in range(4):
            g.launch()
            got = sim.state_vector()  # synchronizes
            assert np.array_equal(got, eager), f"replay {rep}"
            # eager work between replays: on the captured handle and on another one
            sim.evolve([Instruction(G.H(), [0]), Instruction(G.CNot(), [0, n - 1])])
            other.evolve(ins[: len(ins) // 2])
            assert abs(float(np.sum(other.probabilities())) - 1.0) < 1e-12


def test_graph_without_reset_accumulates():
    n = 12
    rng = np.random.default_rng(7)
    ins = random_circuit(n, 60, rng, kinds=FUSABLE)
    assert _plan_singles(n, L.BSIM_SV, ins) == 0
    _say("graph sv without reset")
    ref = StateVectorSimulator(n, 0)
    sim = StateVectorSimulator(n, 0)
    g = sim.capture(ins, reset=False)
    for rep in range(3):
        ref.evolve(ins)
        g.launch()
        assert np.array_equal(sim.state_vector(), ref.state_vector()), f"replay {rep}"
    g.close()
    g.close()  # idempotent
    with pytest.raises(BsimError):
        g.launch()


def test_graph_density_matrix_with_noise():
    n = 6
    ins = []
    for q in range(n):
        ins.append(Instruction(G.Ry(0.3 + 0.2 * q), [q]))
        ins.append(Instruction(Nz.Depolarizing(0.02), [q]))
    for q in range(n - 1):
        ins.append(Instruction(G.CNot(), [q, q + 1]))
    ins += [Instruction(Nz.PhaseFlip(0.1), [2]), Instruction(Nz.AmplitudeDamping(0.05), [4]),
            Instruction(G.Rz(0.7), [5])]
    assert _plan_singles(n, L.BSIM_DM, ins) == 0
    _say("graph dm noisy physical")
    eager = DensityMatrixSimulator(n, 0, noise_mode="physical").evolve(ins).density_matrix()
    np.testing.assert_allclose(eager.ravel(order="F"), O.evolve_dm(n, ins, "physical"), atol=1e-12)
    sim = DensityMatrixSimulator(n, 0, noise_mode="physical")
    with sim.capture(ins) as g:
        for rep in range(3):
            g.launch()
            assert np.array_equal(sim.density_matrix(), eager), f"replay {rep}"
            sim.apply_gate(G.X(), 1)


def test_graph_density_matrix_reference_noise():
    n = 6
    ins = W.brickwall(n, 3, seed=8, noise=True)  # reference Kraus ops inside the tile passes
    assert _plan_singles(n, L.BSIM_DM, ins, "reference") == 0
    _say("graph dm noisy reference")
    eager = DensityMatrixSimulator(n, 0).evolve(ins).density_matrix()
    np.testing.assert_allclose(eager.ravel(order="F"), O.evolve_dm(n, ins, "reference"), atol=1e-12)
    sim = DensityMatrixSimulator(n, 0)
    with sim.capture(ins) as g:
        for rep in range(3):
            g.launch()
            assert np.array_equal(sim.density_matrix(), eager), f"replay {rep}"
            sim.apply_noise(Nz.Depolarizing(0.2), 3)


def test_graph_rejects_unfusable_ops():
    n = 8
    sim = StateVectorSimulator(n, 0)
    _say("graph rejection")
    with pytest.raises(BsimError, match="fusable"):
        sim.capture([Instruction(G.H(), [0]),
                     Instruction(G.Unitary(random_unitary(3, np.random.default_rng(1))), [0, 1, 2])])
    dm = DensityMatrixSimulator(4, 0)  # a 3-qubit reference-mode Kraus channel runs in its own kernel
    with pytest.raises(BsimError, match="fusable"):
        dm.capture([Instruction(Nz.Kraus([np.eye(8)]), [0, 1, 3])])
    # the handle still works after a refused capture
    sim.evolve([Instruction(G.H(), [0])])
    assert abs(sim.probabilities()[0] - 0.5) < 1e-14


def test_autotune_picks_a_valid_config(keep_tile_config):
    n = 16
    ins = W.brickwall(n, 6, seed=3)
    _say("autotune 16q")
    sim = StateVectorSimulator(n, 0)
    r = sim.autotune(ins)
    t, lo = r["tile_bits"], r["low_bits"]
    assert 10 <= t <= 13 and 2 <= lo <= min(6, t - 2) and r["ms"] > 0
    assert r["sweep_bits"] in (3, 4) and r["fusion_limit"] in (0, 48, 96, 192, 384)
    cur_t, cur_lo, cur_rb, cur_fl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L.call("bsim_get_tile_config", ctypes.byref(cur_t), ctypes.byref(cur_lo))
    L.call("bsim_set_sweep_bits", 0, ctypes.byref(cur_rb))
    L.call("bsim_set_fusion_limit", r["fusion_limit"], ctypes.byref(cur_fl))
    assert (cur_t.value, cur_lo.value, cur_rb.value, cur_fl.value) == (t, lo, r["sweep_bits"], r["fusion_limit"])
    sim.reinit(n, 0)
    got = sim.evolve(ins).state_vector()
    np.testing.assert_allclose(got, O.evolve_sv(n, ins), atol=1e-12)


def test_autotune_small_state_keeps_config(keep_tile_config):
    n = 8  # below the smallest candidate tile: nothing to try
    t0, lo0 = ctypes.c_int(), ctypes.c_int()
    L.call("bsim_get_tile_config", ctypes.byref(t0), ctypes.byref(lo0))
    r = StateVectorSimulator(n, 0).autotune(W.brickwall(n, 2, seed=1))
    assert (r["tile_bits"], r["low_bits"], r["ms"]) == (t0.value, lo0.value, 0.0)


def test_fusion_limits_match_oracle(keep_tile_config):
    """Every fusion width the autotuner may pick gives the oracle's state."""
    n = 14
    ins = W.brickwall(n, 8, seed=4)
    exp = O.evolve_sv(n, ins)
    for lim in (48, 96, 192, 384, 0):
        _say(f"fusion limit {lim}")
        L.call("bsim_set_fusion_limit", lim, None)
        sim = StateVectorSimulator(n, 0)
        np.testing.assert_allclose(sim.evolve(ins).state_vector(), exp, atol=1e-12)


def test_graph_after_reinit_is_refused():
    """A graph outlives a re-created state: launching it fails cleanly (the
    state was destroyed, even if the new one sits at the same address) and
    closing it still works; the new state is untouched."""
    n = 10
    ins = W.brickwall(n, 3, seed=2)
    sim = StateVectorSimulator(n, 0)
    g = sim.capture(ins)
    g.launch()
    ref = sim.state_vector()
    np.testing.assert_allclose(ref, O.evolve_sv(n, ins), atol=1e-12)
    _say("graph: reinit with two more qubits, then launch")
    sim.reinit(n + 2, 0)
    with pytest.raises(BsimError):
        g.launch()
    g.close()
    p = sim.probabilities()
    assert p[0] == 1.0 and p.sum() == 1.0
    # the same through the C ABI: destroy the state, keep the graph
    lib = L.load()
    h = ctypes.c_void_p()
    L.call("bsim_state_create", n, L.BSIM_SV, ctypes.byref(h))
    ol = lower_instructions(ins, L.BSIM_SV)
    graph = ctypes.c_void_p()
    L.call("bsim_graph_create", h, ol.ptr, ol.n, 1, ctypes.byref(graph))
    L.call("bsim_graph_launch", graph)
    L.call("bsim_state_destroy", h)
    assert lib.bsim_graph_launch(graph) == L.BSIM_EINVAL
    assert "destroyed" in L.last_error()
    assert lib.bsim_graph_destroy(graph) == 0


def test_autotune_merge_width(keep_tile_config):
    """A 21-qubit circuit of 3- and 4-target unitaries: the autotuner also
    searches the merged block width (off, 4, 5 bits here), and every width
    gives the oracle's state."""
    n = 21
    rng = np.random.default_rng(17)
    ins = []
    for _ in range(5):
        k = int(rng.integers(3, 5))
        q = [int(x) for x in rng.choice(n, k, replace=False)]
        ins.append(Instruction(G.Unitary(random_unitary(k, rng)), q))
        ins.append(Instruction(G.Rx(0.4), [q[0]]))
    exp = O.evolve_sv(n, ins)
    _say("autotune merge width 21q")
    r = StateVectorSimulator(n, 0).autotune(ins)
    assert r["merge_width"] in (0, 4, 5, 6) and r["ms"] > 0
    for mw in (0, 4, 5, 6):
        L.call("bsim_set_merge_width", mw, None)
        sim = StateVectorSimulator(n, 0)
        sim.evolve(ins)
        np.testing.assert_allclose(sim.state_vector(), exp, atol=1e-10, err_msg=f"merge width {mw}")
