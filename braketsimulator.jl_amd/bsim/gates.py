"""Gate types and their matrices, as BraketSimulator.jl defines them.

Matrices follow ``matrix_rep`` / ``matrix_rep_raw`` of src/gate_kernels.jl
(:225-478) and src/custom_gates.jl (:41-360, :431-457), including the
``pow_exponent`` rules.  Each gate also knows which reference kernel it
reaches (src/gate_kernels.jl:675-749) and lowers itself to libbsim ops with
that kernel's index convention:

* 1 target: the 2x2 kernel (:489-537);
* 2 targets, not controlled: the 4x4 kernel for ``Vector`` states
  (:540-576), where the first target is bit 0 of the matrix index;
* CNot/CY/CZ/CV/CSwap/CCNot and ``Control``: the controlled kernels
  (:579-670) with the target gate's matrix;
* 3+ targets: the k-qubit fallback (:753-775), whose matrix bit j is the
  j-th smallest endian target -- the given order is ignored;
* DoubleExcitation: its own kernel (src/custom_gates.jl:390-416).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SQ2 = 1.0 / math.sqrt(2.0)

I2 = np.eye(2, dtype=np.complex128)
PX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
PY = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
PZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)


def _mpow(m: np.ndarray, p: float) -> np.ndarray:
    """Principal matrix power (unitaries are normal: eigen-decomposition)."""
    if float(p).is_integer():
        return np.linalg.matrix_power(m, int(p)) if p >= 0 else np.linalg.matrix_power(np.linalg.inv(m), int(-p))
    w, v = np.linalg.eig(m)
    return (v * (w.astype(np.complex128) ** p)) @ np.linalg.inv(v)


def _diag(*d) -> np.ndarray:
    return np.diag(np.array(d, dtype=np.complex128))


@dataclass
class Gate:
    """Base gate: ``angles`` (possibly empty) and ``pow_exponent``."""

    angles: tuple = ()
    pow_exponent: float = 1.0
    n_qubits: int = 1

    name = "Gate"
    controlled_by = 0  # number of leading control qubits for CNot-like gates

    # --- matrices ---------------------------------------------------------
    def raw(self, *angles) -> np.ndarray:  # matrix_rep_raw
        raise NotImplementedError

    def matrix_rep(self) -> np.ndarray:
        """src/gate_kernels.jl:471-478 generic rule."""
        n = self.pow_exponent
        if n == 0:
            return np.eye(2 ** self.n_qubits, dtype=np.complex128)
        if n == 1:
            return self.raw(*self.angles)
        return _mpow(self.raw(*self.angles), n)

    def target_gate(self):
        """Gate applied to the target qubits by a controlled kernel."""
        return None

    @property
    def parametrized(self) -> bool:
        return len(self.angles) > 0

    def with_angles(self, angles):
        g = type(self).__new__(type(self))
        g.__dict__.update(self.__dict__)
        g.angles = tuple(angles)
        return g

    def __repr__(self):
        a = f"{self.angles}" if self.angles else ""
        p = f"^{self.pow_exponent}" if self.pow_exponent != 1 else ""
        return f"{self.name}{a}{p}"


def _gate(name, nq, raw, angles=0, rule=None, target=None, ncontrols=0):
    """Define a gate class.  ``rule`` overrides matrix_rep; ``target`` is
    the target-gate factory of a controlled gate."""

    def __init__(self, *angles_, pow_exponent=1.0):
        if len(angles_) != angles:
            raise ValueError(f"{name} takes {angles} angle(s)")
        Gate.__init__(self, tuple(float(a) for a in angles_), float(pow_exponent), nq)

    ns = {"__init__": __init__, "name": name, "controlled_by": ncontrols, "raw": lambda self, *a: raw(*a)}
    if rule is not None:
        ns["matrix_rep"] = rule
    if target is not None:
        ns["target_gate"] = lambda self: target(self)
    return type(name, (Gate,), ns)


# ---- matrix_rep rules (src/gate_kernels.jl:242-398) --------------------

def _rule_parity(base_raw_self):  # Y, H, Swap, CNot..., GPi, ECR: (:243-252)
    def rule(self):
        n = self.pow_exponent
        if float(n).is_integer() and int(n) % 2 == 0:
            return np.eye(2 ** self.n_qubits, dtype=np.complex128)
        if float(n).is_integer() and int(n) % 2 == 1:
            return self.raw(*self.angles)
        return _mpow(self.raw(*self.angles), n)
    return rule


def _rule_sqrt(half_raw):  # X (V at 0.5), Z (S at 0.5): (:254-264)
    def rule(self):
        n = self.pow_exponent
        if float(n).is_integer() and int(n) % 2 == 0:
            return np.eye(2, dtype=np.complex128)
        if float(n).is_integer() and int(n) % 2 == 1:
            return self.raw()
        if n == 0.5:
            return half_raw()
        return _mpow(self.raw(), n)
    return rule


def _rule_angled(self):  # Rx, Ry, Rz, PhaseShift (:291-301)
    n = self.pow_exponent
    th = self.angles[0]
    if n == 0:
        return np.eye(2, dtype=np.complex128)
    if n == 1:
        return self.raw(th)
    if float(n).is_integer():
        return self.raw(th * n)
    return _mpow(self.raw(th), n)


def _rule_angled_diag2(self):  # CPhaseShift*, ZZ (:303-311)
    n = self.pow_exponent
    th = self.angles[0]
    if n == 0:
        return np.eye(4, dtype=np.complex128)
    return self.raw(th * n) if n != 1 else self.raw(th)


def _rule_angled4(self):  # XX, YY, XY, SingleExcitation*, FermionicSWAP (:313-322)
    n = self.pow_exponent
    th = self.angles[0]
    if n == 0:
        return np.eye(4, dtype=np.complex128)
    if n == 1:
        return self.raw(th)
    if float(n).is_integer():
        return self.raw(th * n)
    return _mpow(self.raw(th), n)


def _rule_scaled(self):  # DoubleExcitation*, MultiRZ (:324-332)
    n = self.pow_exponent
    th = self.angles[0]
    if n == 0:
        return np.eye(2 ** self.n_qubits, dtype=np.complex128)
    return self.raw(th) if n == 1 else self.raw(th * n)


def _rule_mod4(g1, g2, g3):  # V/Vi, S/Si (:350-371)
    def rule(self):
        n = self.pow_exponent
        if float(n).is_integer():
            r = int(n) % 4
            return [np.eye(2, dtype=np.complex128), g1(), g2(), g3()][r]
        return _mpow(self.raw(), n)
    return rule


# ---- raw matrices ------------------------------------------------------

def _u(th, ph, lam):  # :373-376
    c, s = math.cos(th / 2), math.sin(th / 2)
    return np.array([[c, -np.exp(1j * lam) * s], [np.exp(1j * ph) * s, np.exp(1j * (lam + ph)) * c]])


def _rx(t):
    s, c = math.sin(t / 2.0), math.cos(t / 2.0)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


def _ry(t):
    s, c = math.sin(t / 2.0), math.cos(t / 2.0)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def _rz(t):
    s, c = math.sin(t / 2.0), math.cos(t / 2.0)
    return np.array([[c - 1j * s, 0], [0, c + 1j * s]], dtype=np.complex128)


def _phase(t):
    return _diag(1.0, np.exp(1j * t))


def _swap_like(f):
    return np.array([[1, 0, 0, 0], [0, 0, f, 0], [0, f, 0, 0], [0, 0, 0, 1]], dtype=np.complex128)


def _xx(t):
    s, c = math.sin(t / 2.0), math.cos(t / 2.0)
    return np.array([[c, 0, 0, -1j * s], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [-1j * s, 0, 0, c]])


def _yy(t):
    s, c = math.sin(t / 2.0), math.cos(t / 2.0)
    return np.array([[c, 0, 0, 1j * s], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [1j * s, 0, 0, c]])


def _xy(t):
    s, c = math.sin(t / 2.0), math.cos(t / 2.0)
    return np.array([[1, 0, 0, 0], [0, c, 1j * s, 0], [0, 1j * s, c, 0], [0, 0, 0, 1]], dtype=np.complex128)


def _zz(t):
    e = np.exp(-1j * t / 2.0)
    return _diag(e, np.conj(e), np.conj(e), e)


def _ms(a, b, c):  # :416-425
    cc, ss = math.cos(c / 2.0), math.sin(c / 2)
    return np.array([
        [cc, 0, 0, -1j * np.exp(-1j * (a + b)) * ss],
        [0, cc, -1j * np.exp(-1j * (a - b)) * ss, 0],
        [0, -1j * np.exp(1j * (a - b)) * ss, cc, 0],
        [-1j * np.exp(1j * (a + b)) * ss, 0, 0, cc],
    ])


def _gpi(p):
    return np.array([[0, np.exp(-1j * p)], [np.exp(1j * p), 0]])


def _gpi2(p):
    return SQ2 * np.array([[1.0, -1j * np.exp(-1j * p)], [-1j * np.exp(1j * p), 1.0]])


def _prx(t, p):  # :402-407
    c, s = math.cos(t / 2.0), math.sin(t / 2.0)
    return np.array([[c, -1j * np.exp(-1j * p) * s], [-1j * np.exp(1j * p) * s, c]])


def _single_exc(t, phase=0.0):  # custom_gates.jl SingleExcitation(+/-)
    s, c = math.sin(t / 2.0), math.cos(t / 2.0)
    e = np.exp(1j * phase * t / 2.0)
    return np.array([[e, 0, 0, 0], [0, c, s, 0], [0, -s, c, 0], [0, 0, 0, e]], dtype=np.complex128)


def _fswap(t):
    s, c = math.sin(t / 2.0), math.cos(t / 2.0)
    e2 = np.exp(1j * t / 2.0)
    ie2 = 1j * e2
    return np.array([[1, 0, 0, 0], [0, e2 * c, -ie2 * s, 0], [0, -ie2 * s, e2 * c, 0], [0, 0, 0, np.exp(1j * t)]])


def _double_exc(t, phase=0.0):  # custom_gates.jl:41-49 (and Plus/Minus)
    s, c = math.sin(t / 2.0), math.cos(t / 2.0)
    m = np.eye(16, dtype=np.complex128) * np.exp(1j * phase * t / 2.0)
    m[3, 3] = c
    m[12, 12] = c
    m[3, 12] = -s
    m[12, 3] = s
    return m


_H = SQ2 * np.array([[1, 1], [1, -1]], dtype=np.complex128)
_V = np.array([[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]])
_Vi = np.array([[0.5 - 0.5j, 0.5 + 0.5j], [0.5 + 0.5j, 0.5 - 0.5j]])
_S = _diag(1.0, 1j)
_Si = _diag(1.0, -1j)
_T = _diag(1.0, np.exp(1j * math.pi / 4.0))
_Ti = _diag(1.0, np.exp(-1j * math.pi / 4.0))
_ECR = SQ2 * np.array([[0, 1, 0, 1j], [1, 0, -1j, 0], [0, 1j, 0, 1], [-1j, 0, 1, 0]])
_CNOT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=np.complex128)
_CY = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]])
_CZ = _diag(1, 1, 1, -1)
_CV = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0.5 + 0.5j, 0.5 - 0.5j], [0, 0, 0.5 - 0.5j, 0.5 + 0.5j]])
_CCNOT = np.eye(8, dtype=np.complex128)[[0, 1, 2, 3, 4, 5, 7, 6]]
_CSWAP = np.eye(8, dtype=np.complex128)[[0, 1, 2, 3, 4, 6, 5, 7]]

Identity = _gate("I", 1, lambda: np.eye(2, dtype=np.complex128))
I = Identity
H = _gate("H", 1, lambda: _H.copy(), rule=_rule_parity(None))
X = _gate("X", 1, lambda: PX.copy(), rule=_rule_sqrt(lambda: _V.copy()))
Y = _gate("Y", 1, lambda: PY.copy(), rule=_rule_parity(None))
Z = _gate("Z", 1, lambda: PZ.copy(), rule=_rule_sqrt(lambda: _S.copy()))
V = _gate("V", 1, lambda: _V.copy(), rule=_rule_mod4(lambda: _V.copy(), lambda: PX.copy(), lambda: _Vi.copy()))
Vi = _gate("Vi", 1, lambda: _Vi.copy(), rule=_rule_mod4(lambda: _Vi.copy(), lambda: PX.copy(), lambda: _V.copy()))
S = _gate("S", 1, lambda: _S.copy(), rule=_rule_mod4(lambda: _S.copy(), lambda: PZ.copy(), lambda: _Si.copy()))
Si = _gate("Si", 1, lambda: _Si.copy(), rule=_rule_mod4(lambda: _Si.copy(), lambda: PZ.copy(), lambda: _S.copy()))


def _rule_t(sign):
    def rule(self):  # :385-398
        n = self.pow_exponent
        if n == 0:
            return np.eye(2, dtype=np.complex128)
        if n == 1:
            return self.raw()
        if n == -1:
            return (_Ti if sign > 0 else _T).copy()
        return _phase(sign * n * math.pi / 4)
    return rule


T = _gate("T", 1, lambda: _T.copy(), rule=_rule_t(1))
Ti = _gate("Ti", 1, lambda: _Ti.copy(), rule=_rule_t(-1))
Rx = _gate("Rx", 1, _rx, 1, rule=_rule_angled)
Ry = _gate("Ry", 1, _ry, 1, rule=_rule_angled)
Rz = _gate("Rz", 1, _rz, 1, rule=_rule_angled)
PhaseShift = _gate("PhaseShift", 1, _phase, 1, rule=_rule_angled)


def _rule_u(self):  # :378-384
    n = self.pow_exponent
    if n == 0:
        return np.eye(2, dtype=np.complex128)
    if n == 1:
        return self.raw(*self.angles)
    if n == -1:
        th, ph, lam = self.angles
        return self.raw(-th, -lam, -ph)
    return _mpow(self.raw(*self.angles), n)


U = _gate("U", 1, _u, 3, rule=_rule_u)
GPi = _gate("GPi", 1, _gpi, 1, rule=_rule_parity(None))
GPi2 = _gate("GPi2", 1, _gpi2, 1)


def _rule_prx(self):  # :283-289
    n = self.pow_exponent
    if n == 0:
        return np.eye(2, dtype=np.complex128)
    if n == 1:
        return self.raw(*self.angles)
    if n == -1:
        return self.raw(-self.angles[0], self.angles[1])
    return _mpow(self.raw(*self.angles), n)


PRx = _gate("PRx", 1, _prx, 2, rule=_rule_prx)


def _rule_ms(self):  # :342-348
    n = self.pow_exponent
    if n == 0:
        return np.eye(4, dtype=np.complex128)
    if n == 1:
        return self.raw(*self.angles)
    if n == -1:
        a, b, c = self.angles
        return self.raw(a + math.pi, b, c)
    return _mpow(self.raw(*self.angles), n)


MS = _gate("MS", 2, _ms, 3, rule=_rule_ms)
Swap = _gate("Swap", 2, lambda: _swap_like(1.0), rule=_rule_parity(None))


def _rule_iswap(self):  # :266-272
    n = self.pow_exponent
    if n == 0:
        return np.eye(4, dtype=np.complex128)
    if n == 1:
        return self.raw()
    if n == -1:
        return _swap_like(-1j)
    return _mpow(self.raw(), n)


ISwap = _gate("ISwap", 2, lambda: _swap_like(1j), rule=_rule_iswap)


def _rule_pswap(self):  # :274-281
    n = self.pow_exponent
    th = self.angles[0]
    if n == 0:
        return np.eye(4, dtype=np.complex128)
    if n == 1:
        return self.raw(th)
    if float(n).is_integer() and int(n) % 2 == 1:
        return self.raw(th * n)
    if float(n).is_integer():
        e = np.exp(1j * th * n)
        return _diag(1.0, e, e, 1.0)
    return _mpow(self.raw(th), n)


PSwap = _gate("PSwap", 2, lambda t: _swap_like(np.exp(1j * t)), 1, rule=_rule_pswap)
XX = _gate("XX", 2, _xx, 1, rule=_rule_angled4)
YY = _gate("YY", 2, _yy, 1, rule=_rule_angled4)
XY = _gate("XY", 2, _xy, 1, rule=_rule_angled4)
ZZ = _gate("ZZ", 2, _zz, 1, rule=_rule_angled_diag2)
CPhaseShift = _gate("CPhaseShift", 2, lambda t: _diag(1, 1, 1, np.exp(1j * t)), 1, rule=_rule_angled_diag2)
CPhaseShift00 = _gate("CPhaseShift00", 2, lambda t: _diag(np.exp(1j * t), 1, 1, 1), 1, rule=_rule_angled_diag2)
CPhaseShift01 = _gate("CPhaseShift01", 2, lambda t: _diag(1, 1, np.exp(1j * t), 1), 1, rule=_rule_angled_diag2)
CPhaseShift10 = _gate("CPhaseShift10", 2, lambda t: _diag(1, np.exp(1j * t), 1, 1), 1, rule=_rule_angled_diag2)
ECR = _gate("ECR", 2, lambda: _ECR.copy(), rule=_rule_parity(None))
SingleExcitation = _gate("SingleExcitation", 2, lambda t: _single_exc(t), 1, rule=_rule_angled4)
SingleExcitationPlus = _gate("SingleExcitationPlus", 2, lambda t: _single_exc(t, 1.0), 1, rule=_rule_angled4)
SingleExcitationMinus = _gate("SingleExcitationMinus", 2, lambda t: _single_exc(t, -1.0), 1, rule=_rule_angled4)
FermionicSWAP = _gate("FermionicSWAP", 2, _fswap, 1, rule=_rule_angled4)
DoubleExcitation = _gate("DoubleExcitation", 4, lambda t: _double_exc(t), 1, rule=_rule_scaled)
DoubleExcitationPlus = _gate("DoubleExcitationPlus", 4, lambda t: _double_exc(t, 1.0), 1, rule=_rule_scaled)
DoubleExcitationMinus = _gate("DoubleExcitationMinus", 4, lambda t: _double_exc(t, -1.0), 1, rule=_rule_scaled)

# controlled built-ins: (control gate, target gate, n_controls)  (:720-732)
CNot = _gate("CNot", 2, lambda: _CNOT.copy(), rule=_rule_parity(None), target=lambda g: X(pow_exponent=g.pow_exponent), ncontrols=1)
CY = _gate("CY", 2, lambda: _CY.copy(), rule=_rule_parity(None), target=lambda g: Y(pow_exponent=g.pow_exponent), ncontrols=1)
CZ = _gate("CZ", 2, lambda: _CZ.copy(), rule=_rule_parity(None), target=lambda g: Z(pow_exponent=g.pow_exponent), ncontrols=1)
CV = _gate("CV", 2, lambda: _CV.copy(), target=lambda g: V(pow_exponent=g.pow_exponent), ncontrols=1)
CSwap = _gate("CSwap", 3, lambda: _CSWAP.copy(), rule=_rule_parity(None), target=lambda g: Swap(pow_exponent=g.pow_exponent), ncontrols=1)
CCNot = _gate("CCNot", 3, lambda: _CCNOT.copy(), rule=_rule_parity(None), target=lambda g: X(pow_exponent=g.pow_exponent), ncontrols=2)


class Unitary(Gate):
    """Arbitrary unitary (src/gate_kernels.jl:464, :679)."""

    name = "Unitary"

    def __init__(self, matrix, pow_exponent=1.0):
        m = np.asarray(matrix, dtype=np.complex128)
        k = int(round(math.log2(m.shape[0])))
        Gate.__init__(self, (), float(pow_exponent), k)
        self.matrix = m

    def raw(self):
        return self.matrix.copy()


class MultiRZ(Gate):
    """custom_gates.jl:339-344, :383-385: Rz (1q), ZZ (2q), else a diagonal
    exp(-i theta/2 * eigenvalue) over the Pauli-Z eigenvalues."""

    name = "MultiRZ"

    def __init__(self, angle, pow_exponent=1.0):
        Gate.__init__(self, (float(angle),), float(pow_exponent), 1)

    def diag_on(self, k: int) -> np.ndarray:
        ev = np.array([(-1.0) ** bin(i).count("1") for i in range(2 ** k)])
        return np.exp(-1j * self.angles[0] / 2.0 * ev)


class GPhase(Gate):
    """GPhase{N} (custom_gates.jl:355-362, :386): every amplitude times
    exp(i angle pow)."""

    name = "GPhase"

    def __init__(self, n_qubits: int, angle, pow_exponent=1.0):
        Gate.__init__(self, (float(angle),), float(pow_exponent), int(n_qubits))

    def raw(self, a):
        return np.eye(2 ** self.n_qubits, dtype=np.complex128) * np.exp(1j * a)

    def matrix_rep(self):
        return self.raw(self.angles[0] * self.pow_exponent)


class Control(Gate):
    """Control{G,B} (custom_gates.jl:418-457): B leading control qubits
    with bit values ``bitvals`` on gate ``g``."""

    name = "Control"

    def __init__(self, g: Gate, bitvals, pow_exponent=1.0):
        if isinstance(g, Control):
            bitvals = tuple(g.bitvals) + tuple(bitvals)
            pow_exponent = g.pow_exponent * pow_exponent
            g = g.g
        self.g = g
        self.bitvals = tuple(int(b) for b in bitvals)
        nq = g.n_qubits if isinstance(g, GPhase) else g.n_qubits + len(self.bitvals)
        Gate.__init__(self, g.angles, float(pow_exponent), nq)

    @property
    def angles_(self):
        return self.g.angles

    def with_angles(self, angles):
        return Control(self.g.with_angles(angles), self.bitvals, self.pow_exponent)

    def inner(self) -> Gate:
        """gate.g ^ gate.pow_exponent (:677)."""
        g = self.g.with_angles(self.g.angles)
        g.pow_exponent = self.g.pow_exponent * self.pow_exponent
        return g

    def matrix_rep(self):  # :431-444
        if isinstance(self.g, GPhase):
            # :445-457: a 2^N diagonal (the controls are the phase's own
            # qubits) with the phase at the control index, then the power
            d = np.ones(2 ** self.g.n_qubits, dtype=np.complex128)
            ctrl = sum(b << (b_ix - 1) for b_ix, b in enumerate(self.bitvals, start=1))
            d[ctrl] = np.exp(1j * self.g.angles[0] * self.g.pow_exponent)
            m = np.diag(d)
            return m if self.pow_exponent == 1.0 else _mpow(m, self.pow_exponent)
        inner = self.inner().matrix_rep()
        iq = self.g.n_qubits
        tot = iq + len(self.bitvals)
        full = np.eye(2 ** tot, dtype=np.complex128)
        ctrl = 0
        for b_ix, b in enumerate(self.bitvals, start=1):
            ctrl += b << (b_ix + iq - 1)
        full[ctrl:ctrl + 2 ** iq, ctrl:ctrl + 2 ** iq] = inner
        return full


GATES = {
    c.name: c
    for c in (
        Identity, H, X, Y, Z, V, Vi, S, Si, T, Ti, Rx, Ry, Rz, PhaseShift, U, GPi, GPi2, PRx, MS, Swap, ISwap, PSwap,
        XX, YY, XY, ZZ, CPhaseShift, CPhaseShift00, CPhaseShift01, CPhaseShift10, ECR, SingleExcitation,
        SingleExcitationPlus, SingleExcitationMinus, FermionicSWAP, DoubleExcitation, DoubleExcitationPlus,
        DoubleExcitationMinus, CNot, CY, CZ, CV, CSwap, CCNot,
    )
}


# ---- lowering to kernel calls ------------------------------------------

@dataclass
class KernelCall:
    """One reference-kernel application: ``targets[b]`` is the qubit of
    matrix index bit b, ``controls``/``cvals`` select the subspace."""

    targets: list
    matrix: np.ndarray
    controls: list = field(default_factory=list)
    cvals: list = field(default_factory=list)
    diagonal: bool = False
    double_excitation: float | None = None  # angle for the dedicated kernel


def _sorted_targets(targets):
    """k-qubit fallback: matrix bit j = j-th smallest endian target, i.e.
    the qubits in descending order (src/gate_kernels.jl:759-764)."""
    return sorted(targets, reverse=True)


def _dense_call(m, targets):
    k = len(targets)
    if k == 1:
        return KernelCall([targets[0]], m)
    if k == 2:
        return KernelCall([targets[0], targets[1]], m)  # t1 = bit 0 (:555-563)
    return KernelCall(_sorted_targets(targets), m)


def kernel_call(g: Gate, targets) -> KernelCall:
    """Which kernel the reference dispatches ``apply_gate!(g, sv, targets...)``
    to, and with which matrix (src/gate_kernels.jl:675-749)."""
    targets = [int(t) for t in targets]
    if len(targets) != g.n_qubits and not isinstance(g, (MultiRZ, GPhase)):
        raise ValueError(f"{g!r} acts on {g.n_qubits} qubits, got targets {targets}")
    if isinstance(g, Control):
        if isinstance(g.g, GPhase):  # :734-749
            n = len(targets)
            d = np.ones(2 ** n, dtype=np.complex128)
            idx = 0
            for b_ix, b in enumerate(g.bitvals, start=1):
                idx += b << (b_ix - 1)
            d[idx] = np.exp(1j * g.g.angles[0] * g.g.pow_exponent * g.pow_exponent)
            call = _dense_call(np.diag(d), targets)
            call.matrix = np.diag(call.matrix).copy()
            call.diagonal = True
            return call
        nb = len(g.bitvals)
        inner = g.inner().matrix_rep()
        return KernelCall(list(targets[nb:]), inner, list(targets[:nb]), list(g.bitvals))
    if g.controlled_by:
        nb = g.controlled_by
        tg = g.target_gate()
        return KernelCall(list(targets[nb:]), tg.matrix_rep(), list(targets[:nb]), [1] * nb)
    if isinstance(g, MultiRZ):
        k = len(targets)
        n = g.pow_exponent
        if k == 1:
            return _dense_call(Rz(g.angles[0], pow_exponent=n).matrix_rep(), targets)
        if k == 2:
            return _dense_call(ZZ(g.angles[0], pow_exponent=n).matrix_rep(), targets)
        # k > 2: the reference passes angle/2 without the exponent (custom_gates.jl:385)
        th = g.angles[0]
        d = np.exp(-1j * th / 2.0 * np.array([(-1.0) ** bin(i).count("1") for i in range(2 ** k)]))
        return KernelCall(_sorted_targets(targets), d, diagonal=True)
    if isinstance(g, GPhase):
        k = len(targets)
        f = np.exp(1j * g.angles[0] * g.pow_exponent)
        return KernelCall(list(targets), np.full(2 ** k, f, dtype=np.complex128), diagonal=True)
    if type(g) is DoubleExcitation:
        # the dedicated kernel (custom_gates.jl:390-416) flips (t1,t2) for |1100>
        # and (t3,t4) for |0011>: matrix_rep in kron order with t1 the MSB
        return KernelCall(list(reversed(targets)), g.matrix_rep(), double_excitation=g.angles[0] * g.pow_exponent)
    return _dense_call(g.matrix_rep(), targets)


# Gates exp(-i theta/2 P) with P^2 = 1: the kernel matrix obeys
# M(theta + pi) = -i P M(theta), so dM/dtheta = M(theta + pi) / 2 exactly.
_HALF_ANGLE = ("Rx", "Ry", "Rz", "XX", "YY", "ZZ", "MultiRZ")
# Two-shift rule: with D(s) = (M(theta + s) - M(theta - s)) / 2, a matrix whose
# entries are trigonometric polynomials of theta with frequencies 0, 1/2 and
# 1 has dM/dtheta = A D(pi) + D(pi/2), A = 1/2 - sin(pi/4): on frequency w the
# right side is i (A sin(w pi) + sin(w pi/2)) = i w for w in {0, 1/2, 1}.
_SHIFT_A = 0.5 - math.sqrt(0.5)


def _pow_one(g: Gate) -> bool:
    while True:
        if g.pow_exponent != 1.0:
            return False
        if not isinstance(g, Control):
            return True
        g = g.g


def _fd_derivative(g: Gate, targets, i: int) -> np.ndarray:
    """8th-order central difference in angle i (error ~1e-14 for these
    trigonometric entries)."""
    h = 1e-2
    acc = None
    for k, wk in enumerate((4 / 5, -1 / 5, 4 / 105, -1 / 280), start=1):
        ap = list(g.angles)
        am = list(g.angles)
        ap[i] += k * h
        am[i] -= k * h
        term = wk * (kernel_call(g.with_angles(ap), targets).matrix - kernel_call(g.with_angles(am), targets).matrix)
        acc = term if acc is None else acc + term
    return acc / h


def derivative_matrices(g: Gate, targets, which=None) -> list:
    """d(kernel matrix)/d(angle i) of ``g`` for every angle i in ``which``
    (default: all).  At pow_exponent 1 every angle of the library's gates
    enters with frequencies 0, 1/2 and 1 (half-angle rotations, full-angle
    phases), so the derivative is exact: one shifted matrix for the
    exp(-i theta/2 P) family, the two-shift rule otherwise.  Other exponents
    scale the frequencies: an 8th-order central difference."""
    which = range(len(g.angles)) if which is None else which
    out = []
    exact = _pow_one(g)
    for i in which:
        if not exact:
            out.append(_fd_derivative(g, targets, i))
            continue

        def at(shift):
            a = list(g.angles)
            a[i] += shift
            return kernel_call(g.with_angles(a), targets).matrix

        if type(g).__name__ in _HALF_ANGLE:
            out.append(0.5 * at(math.pi))
        else:
            out.append(0.5 * (_SHIFT_A * (at(math.pi) - at(-math.pi)) + (at(0.5 * math.pi) - at(-0.5 * math.pi))))
    return out
