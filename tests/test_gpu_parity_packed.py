"""Programs carried as compact op-list images (bsim_ops_pack /
bsim_ops_unpack / bsim_evolve_packed) run exactly as the op lists they were
packed from: state vector and density matrix in both evolve modes, the
adjoint gradient, the persistent batch and a captured graph -- bit for bit,
and the state vector also against the oracle."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O
from bsim import _lib as L
from bsim import noises as Nz
from bsim import adjoint as A
from bsim import workloads as W
from bsim.batch import GradientTask, run_batch
from bsim.simulators import DensityMatrixSimulator, Instruction, OpImage, StateVectorSimulator, lower_instructions

from circuits import random_circuit

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", [0, 1])
def test_image_evolve_sv(mode):
    n = 13
    ins = random_circuit(n, 150, np.random.default_rng(20 + mode))
    ol = lower_instructions(ins, L.BSIM_SV)
    im = OpImage(OpImage.pack(ol).tobytes())
    _say(f"image evolve sv mode={mode}")
    a = StateVectorSimulator(n, 0)
    L.call("bsim_evolve", a.handle, ol.ptr, ol.n, mode)
    b = StateVectorSimulator(n, 0)
    L.call("bsim_evolve", b.handle, im.ptr, im.n, mode)
    c = StateVectorSimulator(n, 0)
    L.call("bsim_evolve_packed", c.handle, im.buf.ctypes.data, im.buf.size, mode)
    sa = a.state_vector()
    assert sa.tobytes() == b.state_vector().tobytes() == c.state_vector().tobytes()
    np.testing.assert_allclose(sa, O.evolve_sv(n, ins), atol=1e-10)


@pytest.mark.parametrize("noise_mode", ["reference", "physical"])
def test_image_evolve_dm(noise_mode):
    n = 6
    rng = np.random.default_rng(31)
    ins = random_circuit(n, 40, rng) + [Instruction(Nz.Depolarizing(0.1), [2]),
                                        Instruction(Nz.TwoQubitDepolarizing(0.05), [0, 4]),
                                        Instruction(Nz.PhaseFlip(0.2), [5])]
    ol = lower_instructions(ins, L.BSIM_DM, noise_mode)
    im = OpImage(OpImage.pack(ol).tobytes())
    for mode in (0, 1):
        _say(f"image evolve dm {noise_mode} mode={mode}")
        a = DensityMatrixSimulator(n, 0, noise_mode=noise_mode)
        L.call("bsim_evolve", a.handle, ol.ptr, ol.n, mode)
        b = DensityMatrixSimulator(n, 0, noise_mode=noise_mode)
        L.call("bsim_evolve_packed", b.handle, im.buf.ctypes.data, im.buf.size, mode)
        assert a.density_matrix().tobytes() == b.density_matrix().tobytes()


def test_image_adjoint_batch_and_graph():
    n = 10
    ins, pmap, _ = W.hea(n, 4, seed=9)
    words, coeffs = W.tfim(n)
    ol = lower_instructions(list(ins), L.BSIM_SV, param_map=pmap)
    im = OpImage(OpImage.pack(ol).tobytes())
    _say("image adjoint")
    sim = StateVectorSimulator(n, 0)
    ev_a, g_a = A.adjoint_gradient(sim, ol, pmap, words, coeffs)
    ev_b, g_b = A.adjoint_gradient(sim, im, pmap, words, coeffs)
    assert ev_a == ev_b and g_a.tobytes() == g_b.tobytes()
    _say("image batch")
    bw = lower_instructions(W.brickwall(9, 12, seed=5), L.BSIM_SV)
    bim = OpImage(OpImage.pack(bw).tobytes())
    ra = run_batch([(9, bw), GradientTask(n, ol, pmap, words, coeffs)], [300, 0], [7, 0])
    rb = run_batch([(9, bim), GradientTask(n, im, pmap, words, coeffs)], [300, 0], [7, 0])
    assert ra[0].expect_z.tobytes() == rb[0].expect_z.tobytes()
    assert (ra[0].samples == rb[0].samples).all()
    assert ra[1].grad.tobytes() == rb[1].grad.tobytes()
    _say("image graph")
    g = C.c_void_p()
    sim.reinit(9, 0)
    L.call("bsim_graph_create", sim.handle, bim.ptr, bim.n, 1, C.byref(g))
    try:
        L.call("bsim_graph_launch", g)
        got = sim.state_vector()
    finally:
        L.call("bsim_graph_destroy", g)
    ref = StateVectorSimulator(9, 0)
    L.call("bsim_evolve", ref.handle, bw.ptr, bw.n, 0)
    assert got.tobytes() == ref.state_vector().tobytes()
