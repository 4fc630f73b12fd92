"""Adjoint-method gradients of <H> for Pauli-sum observables.

The reference defines the AdjointGradient result type (src/results.jl:
90-160) but rejects it (src/pragmas.jl:67-68); here it is computed by
libbsim (bsim_adjoint_gradient): one forward evolution, then one backward
sweep that un-applies every gate to |psi> and |lambda> = H|psi> with the
<lambda|dU/dtheta|psi> products fused into the same kernel.

``ExpectationFunction`` is the torch.autograd.Function counterpart of the
ChainRulesCore.rrule a Julia caller would register (forward = <H>, pullback
= the adjoint gradient).  Parameter-shift and central finite differences
through the plain evolution are provided as checks.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from .gates import Gate
from .simulators import Instruction, StateVectorSimulator, lower_instructions


def _words_buf(words, n):
    for w in words:
        if len(w) != n:
            raise ValueError(f"Pauli word {w!r} must have {n} letters")
    return "".join(words).encode()


def expectation_pauli_sum(sim: StateVectorSimulator, words, coeffs) -> float:
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    out = C.c_double()
    L.call("bsim_expect_pauli_sum", sim.handle, len(words), _words_buf(words, sim.qubit_count),
           c.ctypes.data_as(C.POINTER(C.c_double)), C.byref(out))
    return out.value


def with_params(instructions, param_map, theta):
    """Copy of ``instructions`` with angle (i, j) set to theta[param_map[(i, j)]]."""
    out = list(instructions)
    by_ins = {}
    for (i, j), p in param_map.items():
        by_ins.setdefault(i, {})[j] = p
    for i, sub in by_ins.items():
        g: Gate = out[i].operator
        ang = list(g.angles)
        for j, p in sub.items():
            ang[j] = float(theta[p])
        out[i] = Instruction(g.with_angles(ang), out[i].target)
    return out


def adjoint_gradient(sim, instructions, param_map, words, coeffs, n_params=None, checkpoint_every=0, fused=True):
    """(<H>, dH/dtheta) for the circuit from |0...0>.  ``param_map`` maps
    (instruction index, angle index) -> parameter index.  ``instructions``
    may also be an op list lowered with that map (``lower_instructions(...,
    param_map=...)``), which skips the lowering.  On a
    ``DensityMatrixSimulator`` the circuit may hold noise; the forward pass is
    checkpointed every ``checkpoint_every`` ops (0: ceil(sqrt(ops))) and
    recomputed segment by segment.  State vectors: ``fused`` runs the
    backward sweep as tile passes (both vectors in LDS), otherwise one
    kernel per op.  The simulator keeps the final state."""
    n_params = (max(param_map.values()) + 1 if param_map else 0) if n_params is None else n_params
    kind = getattr(sim, "kind", L.BSIM_SV)
    if hasattr(instructions, "ptr"):  # already lowered (with param_map), e.g. reused across calls
        ol = instructions
    else:
        ol = lower_instructions(list(instructions), kind, noise_mode=getattr(sim, "noise_mode", "reference"),
                                param_map=param_map)
    old = C.c_int()
    L.call("bsim_set_checkpoint_interval", int(checkpoint_every), C.byref(old))
    old_f = C.c_int()
    L.call("bsim_set_adjoint_fusion", 1 if fused else 0, C.byref(old_f))
    grad = np.zeros(n_params, dtype=np.float64)
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    ev = C.c_double()
    L.call("bsim_adjoint_gradient", sim.handle, ol.ptr, ol.n, len(words), _words_buf(words, sim.qubit_count),
           c.ctypes.data_as(C.POINTER(C.c_double)), n_params, grad.ctypes.data_as(C.POINTER(C.c_double)),
           C.byref(ev))
    L.call("bsim_set_checkpoint_interval", old.value, None)
    L.call("bsim_set_adjoint_fusion", old_f.value, None)
    return ev.value, grad


def energy(sim, instructions, words, coeffs) -> float:
    sim.reinit(sim.qubit_count, 0)
    sim.evolve(instructions)
    return expectation_pauli_sum(sim, words, coeffs)


_PAULI_ROTATIONS = ("Rx", "Ry", "Rz")


def parameter_shift_gradient(sim, instructions, param_map, theta, words, coeffs):
    """d<H>/dtheta_p = (f(theta + pi/2 e_p) - f(theta - pi/2 e_p)) / 2 for
    Pauli rotations (exact); 2 evolutions per parameter through the plain
    path."""
    theta = np.asarray(theta, dtype=np.float64)
    grad = np.zeros_like(theta)
    for (i, j), p in param_map.items():
        name = instructions[i].operator.name
        if name not in _PAULI_ROTATIONS:
            raise ValueError(f"parameter shift needs a Pauli rotation, got {name}")
    for p in range(theta.size):
        tp = theta.copy()
        tp[p] += math.pi / 2
        tm = theta.copy()
        tm[p] -= math.pi / 2
        grad[p] = 0.5 * (energy(sim, with_params(instructions, param_map, tp), words, coeffs)
                         - energy(sim, with_params(instructions, param_map, tm), words, coeffs))
    return grad


def finite_difference_gradient(sim, instructions, param_map, theta, words, coeffs, h=1e-6):
    theta = np.asarray(theta, dtype=np.float64)
    grad = np.zeros_like(theta)
    for p in range(theta.size):
        tp = theta.copy()
        tp[p] += h
        tm = theta.copy()
        tm[p] -= h
        grad[p] = (energy(sim, with_params(instructions, param_map, tp), words, coeffs)
                   - energy(sim, with_params(instructions, param_map, tm), words, coeffs)) / (2 * h)
    return grad


try:  # optional: torch.autograd.Function wrapper
    import torch

    class ExpectationFunction(torch.autograd.Function):
        """f(theta) = <H>(theta) with the adjoint gradient as its pullback."""

        @staticmethod
        def forward(ctx, theta, sim, instructions, param_map, words, coeffs):
            th = theta.detach().cpu().numpy().astype(np.float64)
            ins = with_params(instructions, param_map, th)
            ev, grad = adjoint_gradient(sim, ins, param_map, words, coeffs, n_params=th.size)
            ctx.save_for_backward(torch.from_numpy(grad))
            return torch.tensor(ev, dtype=torch.float64)

        @staticmethod
        def backward(ctx, g):
            (grad,) = ctx.saved_tensors
            return g * grad, None, None, None, None, None
except ImportError:  # pragma: no cover
    ExpectationFunction = None
