"""Helpers shared by the tests: fixture ops -> Instructions, the GRCS-16
QASM fixture reader, random circuits."""
from __future__ import annotations

import math
import re

import numpy as np

import bsim
from bsim import gates as G
from bsim import noises as Nz
from bsim.simulators import Instruction

_P = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
      "Z": np.diag([1.0, -1.0])}


def _kron_spec(spec: str):
    body = spec.split(":", 1)[1]
    zero = body.endswith("+zero")
    body = body.replace("+zero", "")
    m = np.ones((1, 1), dtype=np.complex128)
    for ch in body:
        m = np.kron(m, _P[ch])
    return m, zero


def _cmatrix(rows):
    return np.array([[complex(x[0], x[1]) for x in r] for r in rows], dtype=np.complex128)


def gate_from_spec(spec):
    """[name, args] -> Gate; "Control" takes [inner spec, bitvals],
    "GPhaseN" takes [n, angle]."""
    name, args = spec
    if name == "Control":
        return G.Control(gate_from_spec(args[0]), args[1])
    if name == "GPhaseN":
        return G.GPhase(int(args[0]), args[1])
    if name == "MultiRZ":
        return G.MultiRZ(*args)
    return G.GATES[name](*args)


def fixture_instruction(op):
    name, args, targets = op
    if name == "GPhase1":
        return Instruction(G.GPhase(1, args[0]), targets)
    if name == "GPhaseN":
        return Instruction(G.GPhase(int(args[0]), args[1]), targets)
    if name == "MultiRZ":
        return Instruction(G.MultiRZ(*args), targets)
    if name == "ControlOf":
        return Instruction(G.Control(gate_from_spec(args[0]), args[1]), targets)
    if name == "UnitaryOf":  # Unitary(Matrix(matrix_rep(gate)))
        return Instruction(G.Unitary(gate_from_spec(args).matrix_rep()), targets)
    if name == "Unitary":
        m = _kron_spec(args)[0] if isinstance(args, str) else _cmatrix(args)
        return Instruction(G.Unitary(m), targets)
    if name == "Kraus":
        if isinstance(args, str):
            m, zero = _kron_spec(args)
            ms = [m] + ([0.0 * np.ones_like(m)] if zero else [])
        else:
            ms = [_cmatrix(a) for a in args]
        return Instruction(Nz.Kraus(ms), targets)
    if name in G.GATES:
        return Instruction(G.GATES[name](*args), targets)
    if name in Nz.NOISES:
        return Instruction(Nz.NOISES[name](*args), targets)
    raise KeyError(name)


def fixture_ops(case, ops=None):
    return [fixture_instruction(o) for o in (case["ops"] if ops is None else ops)]


def fixture_observable(factors):
    """Observable of a closed-form fixture: one factor or a TensorProduct."""
    from bsim import observables as Ob

    def one(f):
        if isinstance(f, str):
            return {"X": Ob.X, "Y": Ob.Y, "Z": Ob.Z, "H": Ob.H, "I": Ob.I}[f]()
        return Ob.HermitianObservable(_cmatrix(f[1]))

    fs = [one(f) for f in factors]
    return fs[0] if len(fs) == 1 else Ob.TensorProduct(fs)


def grcs16(path):
    """Instruction list of the GRCS-16 OpenQASM fixture (h, cz, t, rx(pi/2),
    ry(pi/2); index lists, inclusive ranges and whole-register forms)."""
    ins = []
    n = 16

    def qubits(spec):
        spec = spec.strip()
        if spec == "q":
            return list(range(n))
        inner = spec[spec.index("[") + 1: spec.rindex("]")]
        if inner.startswith("{"):
            return [int(x) for x in inner.strip("{}").split(",")]
        if ":" in inner:
            a, b = inner.split(":")
            return list(range(int(a), int(b) + 1))
        return [int(inner)]

    for line in path.read_text().splitlines():
        line = line.strip().rstrip(";")
        if not line or line.startswith(("OPENQASM", "qubit", "#")):
            continue
        m = re.match(r"(\w+)(\((.*?)\))?\s+(.*)", line)
        g, _, arg, rest = m.groups()
        if g == "cz":
            a, b = [qubits(x)[0] for x in rest.split(",")]
            ins.append(Instruction(G.CZ(), [a, b]))
            continue
        for q in qubits(rest):
            if g == "h":
                ins.append(Instruction(G.H(), [q]))
            elif g == "t":
                ins.append(Instruction(G.T(), [q]))
            elif g == "rx":
                ins.append(Instruction(G.Rx(math.pi / 2), [q]))
            elif g == "ry":
                ins.append(Instruction(G.Ry(math.pi / 2), [q]))
            else:
                raise ValueError(line)
    return ins


ONE_Q = ["H", "X", "Y", "Z", "S", "Si", "T", "Ti", "V", "Vi"]
ONE_Q_ANG = [("Rx", 1), ("Ry", 1), ("Rz", 1), ("PhaseShift", 1), ("U", 3), ("GPi", 1), ("GPi2", 1), ("PRx", 2)]
TWO_Q = ["Swap", "ISwap", "ECR", "CNot", "CY", "CZ", "CV"]
TWO_Q_ANG = [("XX", 1), ("YY", 1), ("ZZ", 1), ("XY", 1), ("PSwap", 1), ("CPhaseShift", 1), ("CPhaseShift00", 1),
             ("CPhaseShift01", 1), ("CPhaseShift10", 1), ("MS", 3), ("SingleExcitation", 1),
             ("SingleExcitationPlus", 1), ("SingleExcitationMinus", 1), ("FermionicSWAP", 1)]


def random_unitary(k, rng):
    a = rng.normal(size=(2 ** k, 2 ** k)) + 1j * rng.normal(size=(2 ** k, 2 ** k))
    q, r = np.linalg.qr(a)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def random_circuit(n, n_ops, rng, kinds=("1q", "2q", "ctrl", "kq", "diag", "pow")):
    """A random mix of every gate family the reference dispatches."""
    ins = []
    for _ in range(n_ops):
        kind = kinds[rng.integers(len(kinds))]
        if kind == "1q" or n < 2:
            if rng.random() < 0.5:
                g = G.GATES[ONE_Q[rng.integers(len(ONE_Q))]]()
            else:
                name, na = ONE_Q_ANG[rng.integers(len(ONE_Q_ANG))]
                g = G.GATES[name](*rng.uniform(-math.pi, math.pi, na))
            ins.append(Instruction(g, [int(rng.integers(n))]))
        elif kind == "2q":
            t = rng.choice(n, 2, replace=False)
            if rng.random() < 0.4:
                g = G.GATES[TWO_Q[rng.integers(len(TWO_Q))]]()
            elif rng.random() < 0.5:
                name, na = TWO_Q_ANG[rng.integers(len(TWO_Q_ANG))]
                g = G.GATES[name](*rng.uniform(-math.pi, math.pi, na))
            else:
                g = G.Unitary(random_unitary(2, rng))
            ins.append(Instruction(g, [int(x) for x in t]))
        elif kind == "ctrl" and n >= 3:
            t = [int(x) for x in rng.choice(n, 3, replace=False)]
            r = rng.random()
            if r < 0.25:
                ins.append(Instruction(G.CCNot(), t))
            elif r < 0.5:
                ins.append(Instruction(G.CSwap(), t))
            elif r < 0.75:
                g = G.Control(G.Rx(float(rng.uniform(-3, 3))), (int(rng.integers(2)),))
                ins.append(Instruction(g, t[:2]))
            else:
                g = G.Control(G.Ry(float(rng.uniform(-3, 3))), (int(rng.integers(2)), int(rng.integers(2))))
                ins.append(Instruction(g, t))
        elif kind == "kq" and n >= 3:
            k = int(rng.integers(3, min(n, 5) + 1))
            t = [int(x) for x in rng.choice(n, k, replace=False)]
            ins.append(Instruction(G.Unitary(random_unitary(k, rng)), t))
        elif kind == "diag" and n >= 3:
            k = int(rng.integers(1, min(n, 4) + 1))
            t = [int(x) for x in rng.choice(n, k, replace=False)]
            if rng.random() < 0.5:
                ins.append(Instruction(G.MultiRZ(float(rng.uniform(-3, 3))), t))
            elif rng.random() < 0.5:
                bits = tuple(int(x) for x in rng.integers(0, 2, k))
                ins.append(Instruction(G.Control(G.GPhase(k, float(rng.uniform(-3, 3))), bits), t))
            else:
                ins.append(Instruction(G.GPhase(k, float(rng.uniform(-3, 3))), t))
        elif kind == "pow":
            name = ["X", "H", "S", "T", "V", "Rx", "U", "ISwap"][rng.integers(8)]
            p = float(rng.choice([-1.0, 0.5, 2.0, 3.0, 0.3]))
            if name in ("Rx",):
                g = G.Rx(float(rng.uniform(-3, 3)), pow_exponent=p)
            elif name == "U":
                g = G.U(*rng.uniform(-3, 3, 3), pow_exponent=p)
            else:
                g = G.GATES[name](pow_exponent=p)
            k = g.n_qubits
            t = [int(x) for x in rng.choice(n, k, replace=False)]
            ins.append(Instruction(g, t))
        else:
            ins.append(Instruction(G.H(), [int(rng.integers(n))]))
    return ins


def random_noise(n, rng):
    r = rng.random()
    q = int(rng.integers(n))
    if r < 0.2:
        return Instruction(Nz.Depolarizing(float(rng.uniform(0, 0.3))), [q])
    if r < 0.35:
        return Instruction(Nz.AmplitudeDamping(float(rng.uniform(0, 0.3))), [q])
    if r < 0.5:
        return Instruction(Nz.PhaseFlip(float(rng.uniform(0, 0.3))), [q])
    if r < 0.6:
        return Instruction(Nz.GeneralizedAmplitudeDamping(float(rng.uniform(0, 0.3)), float(rng.uniform(0, 1))), [q])
    if r < 0.7:
        return Instruction(Nz.PhaseDamping(float(rng.uniform(0, 0.3))), [q])
    if r < 0.8 and n >= 2:
        t = [int(x) for x in rng.choice(n, 2, replace=False)]
        return Instruction(Nz.TwoQubitDepolarizing(float(rng.uniform(0, 0.2))), t)
    if r < 0.9 and n >= 2:
        t = [int(x) for x in rng.choice(n, 2, replace=False)]
        return Instruction(Nz.TwoQubitDephasing(float(rng.uniform(0, 0.2))), t)
    k1 = random_unitary(1, rng) * math.sqrt(0.7)
    k2 = random_unitary(1, rng) * math.sqrt(0.3)
    return Instruction(Nz.Kraus([k1, k2]), [q])


def chain_circuit(n, layers, seed, n_hot=1):
    """Register-chain workload of the fused planner: ``n_hot`` qubits get a
    U gate per layer, each followed by a CZ to a random other qubit (a Z on
    the register that depends on a thread bit), plus a CZ between two other
    qubits (a per-thread sign) and now and then a Z on a hot qubit (a
    uniform Z).  chain_circuit(13, 9, 0, 2) plans into every chain form:
    chains of more than four matrices, per-thread signs on a long chain and
    sign ops left at the end of a sweep."""
    rng = np.random.default_rng(seed)
    ins = [Instruction(G.H(), [q]) for q in range(n)]
    hot = list(range(n - n_hot, n))
    cold = list(range(0, n - n_hot))
    for _ in range(layers):
        for q in hot:
            ins.append(Instruction(G.U(*rng.uniform(-3, 3, 3)), [q]))
            ins.append(Instruction(G.CZ(), [q, int(rng.choice(cold))]))
        a, b = rng.choice(cold, 2, replace=False)
        ins.append(Instruction(G.CZ(), [int(a), int(b)]))
        if rng.random() < 0.3:
            ins.append(Instruction(G.Z(), [int(rng.choice(hot))]))
    return ins


def dm_sparse_noise(n, depth, seed):
    """A density-matrix brick-wall with two noise channels in the middle:
    the noise-free passes around them plan into register chains."""
    from bsim import workloads as W

    ins = W.brickwall(n, depth, seed=seed)
    ins.append(Instruction(Nz.Depolarizing(0.05), [0]))
    ins.append(Instruction(Nz.AmplitudeDamping(0.1), [n - 1]))
    return ins + W.brickwall(n, 2, seed=seed + 1)


def one_qubit_runs(n, layers, seed):
    """Parametrized circuit whose backward list holds every shape of
    one-qubit run the adjoint planner merges or leaves alone, per layer and
    qubit q (shape q % 7): RZ RY (Z product before the merged op, Y after),
    RX RX (same axis: both products after), RY H RZ (a parameter-free op
    inside the run), RZ RY RX (three rotations: the third stays apart),
    PhaseShift RY (a payload of no Pauli form: no merge), U (three products)
    and RY Control(RZ) (a controlled op ends the run); then CZ on (q, q+1)
    for even q of the layer.  Returns (instructions, param_map, thetas)."""
    rng = np.random.default_rng(seed)
    ins, pmap = [], {}
    p = 0

    def rot(g, qs, k=1):
        nonlocal p
        a = [float(x) for x in rng.uniform(-3, 3, k)]
        ins.append(Instruction(g(*a), qs))
        for j in range(k):
            pmap[(len(ins) - 1, j)] = p
            p += 1

    for layer in range(layers):
        for q in range(n):
            shape = (q + layer) % 7
            if shape == 0:
                rot(G.Rz, [q]); rot(G.Ry, [q])
            elif shape == 1:
                rot(G.Rx, [q]); rot(G.Rx, [q])
            elif shape == 2:
                rot(G.Ry, [q]); ins.append(Instruction(G.H(), [q])); rot(G.Rz, [q])
            elif shape == 3:
                rot(G.Rz, [q]); rot(G.Ry, [q]); rot(G.Rx, [q])
            elif shape == 4:
                rot(G.PhaseShift, [q]); rot(G.Ry, [q])
            elif shape == 5:
                rot(G.U, [q], 3)
            else:
                rot(G.Ry, [q])
                a = float(rng.uniform(-3, 3))
                ins.append(Instruction(G.Control(G.Rz(a), (1,)), [(q + 1) % n, q]))
                pmap[(len(ins) - 1, 0)] = p
                p += 1
        for q in range(layer % 2, n - 1, 2):
            ins.append(Instruction(G.CZ(), [q, q + 1]))
    th = np.zeros(p)
    for (i, j), k in pmap.items():
        th[k] = ins[i].operator.angles[j]
    return ins, pmap, th
