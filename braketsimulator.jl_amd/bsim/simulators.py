"""StateVectorSimulator / DensityMatrixSimulator backed by libbsim.

Mirrors src/sv_simulator.jl and src/dm_simulator.jl: ``reinit``
(reinit!), ``evolve`` (evolve!), ``apply_gate`` (apply_gate!),
``apply_noise`` (apply_noise!), ``expectation``, ``probabilities``,
``state_vector`` / ``density_matrix``, ``apply_observables`` /
``state_with_observables`` and ``samples`` (src/result_types.jl:56-80).
Julia's ``!`` suffix has no Python spelling; the names are otherwise the
reference's.  All amplitudes live on the GPU; host arrays appear only when
a result is read back.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import _lib as L
from . import gates as G
from .gates import Gate, GPhase, KernelCall, Ry, Unitary, H as HGate, kernel_call, derivative_matrices
from .noises import Noise, PhaseFlip
from .observables import H as HObs
from .observables import HermitianObservable, I as IObs, Observable, TensorProduct, X as XObs, Y as YObs
from .observables import Z as ZObs, is_pauli_word, pauli_string


class Instruction:
    """An operator with its target qubits (src/schemas.jl Instruction)."""

    __slots__ = ("operator", "target")

    def __init__(self, operator, target):
        self.operator = operator
        self.target = tuple(int(t) for t in (target if hasattr(target, "__iter__") else [target]))

    def __repr__(self):
        return f"Instruction({self.operator!r}, {list(self.target)})"


class Measure:
    name = "Measure"
    n_qubits = 1


class Barrier:
    name = "Barrier"


class Reset:
    name = "Reset"


class Delay:
    name = "Delay"


NO_OPS = (Measure, Barrier, Reset, Delay)


class Superoperator:
    """A linear map on the density matrix of k <= 3 qubits, given by its
    4^k x 4^k Liouville matrix under column stacking of the physical rho:
    rho'[r, c] = sum S[r + 2^k c, r' + 2^k c'] rho[r', c'], where r and c
    index the targets' block with the first target as the most significant
    bit, as in Braket's Kraus and unitary matrices (U rho U^dagger:
    S = kron(conj(U), U); Kraus K_m: S = sum kron(conj(K_m), K_m)).  One
    fused pass of bsim_dm_apply_superop; the reference applies a DM gate as
    two passes (src/dm_simulator.jl:130-140)."""

    def __init__(self, matrix):
        m = np.asarray(matrix, dtype=np.complex128)
        k = int(round(np.log2(m.shape[0]) / 2))
        if m.ndim != 2 or m.shape != (4 ** k, 4 ** k) or not 1 <= k <= 3:
            raise ValueError("a superoperator is a 4^k x 4^k matrix, 1 <= k <= 3")
        self.matrix = m
        self.qubit_count = k

    @classmethod
    def from_kraus(cls, kraus):
        return cls(sum(np.kron(np.conj(K), K) for K in kraus))

    @classmethod
    def from_unitary(cls, u):
        return cls(np.kron(np.conj(u), u))


# ---- lowering to bsim_op arrays ----------------------------------------

class OpList:
    """An array of bsim_op plus the buffer its matrices live in.  Ops are
    recorded as plain Python values and written into a numpy array laid out
    like the C struct in one go when the array is first used; all matrices
    (column-major complex128) share one buffer, so the pointer fields are
    filled by one vector operation."""

    _DTYPE = None

    @classmethod
    def _dtype(cls):
        if cls._DTYPE is None:
            f = L.Op
            names = ["type", "n_targets", "n_controls", "n_kraus", "targets", "controls", "control_values", "matrix",
                     "n_params", "param_index", "dmatrix"]
            formats = ["<i4", "<i4", "<i4", "<i4", ("<i4", L.MAX_TARGETS), ("<i4", L.MAX_CONTROLS),
                       ("<i4", L.MAX_CONTROLS), "<u8", "<i4", ("<i4", 3), ("<u8", 3)]
            cls._DTYPE = np.dtype({"names": names, "formats": formats,
                                   "offsets": [getattr(f, n).offset for n in names], "itemsize": C.sizeof(f)})
        return cls._DTYPE

    def __init__(self, n: int):
        self.n = 0
        self._ops = []    # (type, targets, controls, cvals, n_kraus, params)
        self._mats = []   # flattened complex128 arrays
        self._where = []  # (op index, slot): slot 0 = matrix, 1 + p = dmatrix[p]
        self._rec = None
        self._arr = None
        # parameters bind() rewrites: Pauli rotations (op index, 0 Rx / 1 Ry /
        # 2 Rz, parameter) and other gates (op index, instruction, {angle: parameter})
        self._rot = []
        self._gen = []
        self._rot_arrays = None
        self._offset = None

    def _mat(self, m, op: int, slot: int):
        a = np.asarray(m, dtype=np.complex128)
        self._mats.append(a.ravel(order="F") if a.ndim == 2 else a.ravel())
        self._where.append((op, slot))

    def add(self, typ, targets, matrix, controls=(), cvals=(), n_kraus=0, params=None, dmats=None):
        i = self.n
        self._ops.append((typ, targets, controls, cvals, n_kraus, params))
        self._mat(matrix, i, 0)
        if params:
            for p, dm in enumerate(dmats):
                self._mat(dm, i, 1 + p)
        self.n += 1
        self._rec = None

    def _finish(self):
        if self._rec is not None:
            return
        rec = np.zeros(max(self.n, 1), dtype=self._dtype())
        if self.n:
            ops = self._ops
            zt, zc = [0] * L.MAX_TARGETS, [0] * L.MAX_CONTROLS
            rec["type"][:self.n] = [o[0] for o in ops]
            rec["n_targets"][:self.n] = [len(o[1]) for o in ops]
            rec["n_controls"][:self.n] = [len(o[2]) for o in ops]
            rec["n_kraus"][:self.n] = [o[4] for o in ops]
            rec["targets"][:self.n] = [list(o[1]) + zt[len(o[1]):] for o in ops]
            if any(o[2] for o in ops):
                rec["controls"][:self.n] = [list(o[2]) + zc[len(o[2]):] for o in ops]
                rec["control_values"][:self.n] = [
                    [int(o[3][k]) if k < len(o[3]) else 1 for k in range(len(o[2]))] + zc[len(o[2]):] for o in ops]
            if any(o[5] for o in ops):
                rec["n_params"][:self.n] = [len(o[5]) if o[5] else 0 for o in ops]
                rec["param_index"][:self.n] = [(list(o[5]) + [0, 0, 0])[:3] if o[5] else [0, 0, 0] for o in ops]
        if self._mats:
            buf = np.concatenate(self._mats)
            sizes = np.fromiter((m.size for m in self._mats), dtype=np.int64, count=len(self._mats))
            self._offset = np.cumsum(sizes) - sizes
            addr = buf.ctypes.data + 16 * self._offset
            where = np.asarray(self._where, dtype=np.int64)
            main = where[:, 1] == 0
            rec["matrix"][where[main, 0]] = addr[main]
            for p in range(3):
                sel = where[:, 1] == 1 + p
                if np.any(sel):
                    rec["dmatrix"][where[sel, 0], p] = addr[sel]
            self._buf = buf
        self._rec = rec
        self._arr = (L.Op * len(rec)).from_buffer(rec)

    @property
    def arr(self):
        self._finish()
        return self._arr

    @staticmethod
    def _call_arrays(call: KernelCall, dmats=None):
        """(op type, matrix, derivative matrices) of a kernel call as add_call
        records them: a diagonal call with derivatives becomes a dense gate."""
        if call.diagonal:
            d = np.asarray(call.matrix, dtype=np.complex128)
            if dmats is not None:
                dmats = [np.diag(np.asarray(x)) if np.ndim(x) == 1 else x for x in dmats]
                return L.OP_GATE, np.diag(d), dmats
            return L.OP_DIAG, d, None
        return L.OP_GATE, call.matrix, dmats

    def add_call(self, call: KernelCall, params=None, dmats=None):
        typ, m, dm = self._call_arrays(call, dmats)
        self.add(typ, call.targets, m, call.controls, call.cvals, params=params, dmats=dm)

    def bind(self, theta) -> "OpList":
        """Rebind the parameters of an op list lowered with a param_map: every
        mapped angle (i, j) becomes theta[param_map[(i, j)]] (as
        adjoint.with_params sets it) and the op's matrix and derivative
        matrices are recomputed in place -- the same numbers a new lowering of
        the rebound instructions gives, without lowering the circuit again.
        Pauli rotations are recomputed together; other gates one by one."""
        self._finish()
        th = np.asarray(theta, dtype=np.float64)
        buf = self._buf if self._mats else None
        if self._rot:
            if self._rot_arrays is None:
                where = {w: int(o) for w, o in zip(self._where, self._offset)}
                r = np.asarray(self._rot, dtype=np.int64)
                self._rot_arrays = (r[:, 1], r[:, 2],
                                    np.asarray([where[(int(i), 0)] for i in r[:, 0]], dtype=np.int64),
                                    np.asarray([where[(int(i), 1)] for i in r[:, 0]], dtype=np.int64))
            code, pidx, m_off, d_off = self._rot_arrays
            t = th[pidx]
            cols = np.arange(4)
            for scale, ang, off in ((1.0, t, m_off), (0.5, t + math.pi, d_off)):
                half = (ang / 2.0).tolist()
                s = np.array([math.sin(x) for x in half])
                c = np.array([math.cos(x) for x in half])
                m = np.zeros((t.size, 4), dtype=np.complex128)
                rx, ry, rz = code == 0, code == 1, code == 2
                m[:, 0] = np.where(rz, c - 1j * s, c)
                m[:, 3] = np.where(rz, c + 1j * s, c)
                m[:, 1] = np.where(rx, -1j * s, np.where(ry, s, 0.0))
                m[:, 2] = np.where(rx, -1j * s, np.where(ry, -s, 0.0))
                buf[off[:, None] + cols] = scale * m if scale != 1.0 else m
        if self._gen:
            where = {w: int(o) for w, o in zip(self._where, self._offset)}
            for op_i, ins, jp in self._gen:
                g = ins.operator
                ang = list(g.angles)
                for j, p in jp.items():
                    ang[j] = float(th[p])
                g2 = g.with_angles(ang)
                js = sorted(jp)
                _, m, dms = self._call_arrays(kernel_call(g2, ins.target),
                                              derivative_matrices(g2, ins.target, js))
                for slot, a in enumerate([m] + list(dms)):
                    a = np.asarray(a, dtype=np.complex128)
                    flat = a.ravel(order="F") if a.ndim == 2 else a.ravel()
                    o = where[(op_i, slot)]
                    buf[o:o + flat.size] = flat
        return self

    @property
    def ptr(self):
        self._finish()
        return C.cast(self._arr, C.POINTER(L.Op))


class OpImage:
    """A compact op-list image (bsim_ops_pack): self-contained bytes that can
    be stored or sent and evolved, captured or batched elsewhere.  ``n`` and
    ``ptr`` (ops pointing into the image, bsim_ops_unpack) as for OpList;
    ``tobytes()`` gives the image."""

    def __init__(self, data):
        self.buf = np.frombuffer(bytes(data), dtype=np.uint8).copy()  # an aligned private copy
        n = C.c_int()
        L.call("bsim_ops_unpack", self.buf.ctypes.data, self.buf.size, None, 0, C.byref(n))
        self._ops = (L.Op * max(n.value, 1))()
        L.call("bsim_ops_unpack", self.buf.ctypes.data, self.buf.size, self._ops, n.value, C.byref(n))
        self.n = n.value

    @classmethod
    def pack(cls, ops) -> "OpImage":
        """The image of an op list (anything with ``n`` and ``ptr``)."""
        size = C.c_size_t()
        L.call("bsim_ops_pack", ops.ptr, ops.n, None, 0, C.byref(size))
        buf = np.zeros(size.value, dtype=np.uint8)
        L.call("bsim_ops_pack", ops.ptr, ops.n, buf.ctypes.data, buf.size, C.byref(size))
        return cls(buf.tobytes())

    @property
    def ptr(self):
        return C.cast(self._ops, C.POINTER(L.Op))

    def tobytes(self) -> bytes:
        return self.buf.tobytes()


# Gates whose kernel call depends only on (class, exponent, targets): their
# calls are reused across lowerings (the matrices are only read).
_FIXED_GATES = tuple(getattr(G, n) for n in ("I", "H", "X", "Y", "Z", "V", "Vi", "S", "Si", "T", "Ti", "CNot", "CY",
                                             "CZ", "CV", "CSwap", "CCNot", "Swap", "ISwap", "ECR"))
_CALL_CACHE: dict = {}
# Pauli rotations at exponent 1: matrix raw(theta) and derivative
# raw(theta + pi) / 2, the numbers kernel_call / derivative_matrices give
_ROTATIONS = {G.Rx: G._rx, G.Ry: G._ry, G.Rz: G._rz}
_ROT_CODE = {G.Rx: 0, G.Ry: 1, G.Rz: 2}


def lower_instructions(instructions, kind: int, noise_mode: str = "reference", param_map=None,
                       fast: bool = True) -> OpList:
    """Instructions -> bsim_op array.  ``param_map`` (adjoint) maps
    (instruction index, angle index) -> global parameter index.  ``fast``
    takes the shortcuts above (the same op array; False: every gate through
    kernel_call / derivative_matrices)."""
    ol = OpList(len(instructions))
    for i, ins in enumerate(instructions):
        op = ins.operator
        if isinstance(op, NO_OPS) or op is None:
            continue
        tp = type(op)
        if fast and tp in _ROTATIONS and op.pow_exponent == 1.0 and len(ins.target) == 1:
            raw, th = _ROTATIONS[tp], op.angles[0]
            params = dmats = None
            if param_map is not None and (i, 0) in param_map:
                params, dmats = [param_map[(i, 0)]], [0.5 * raw(th + math.pi)]
                ol._rot.append((ol.n, _ROT_CODE[tp], params[0]))
            ol.add(L.OP_GATE, [int(ins.target[0])], raw(th), params=params, dmats=dmats)
            continue
        if fast and tp in _FIXED_GATES:
            key = (tp, op.pow_exponent, tuple(int(t) for t in ins.target))
            call = _CALL_CACHE.get(key)
            if call is None:
                if len(_CALL_CACHE) > 65536:
                    _CALL_CACHE.clear()
                call = _CALL_CACHE[key] = kernel_call(op, ins.target)
            ol.add_call(call)
            continue
        if isinstance(op, Gate):
            call = kernel_call(op, ins.target)
            params = dmats = None
            if param_map is not None and op.parametrized:
                idx = [(j, param_map[(i, j)]) for j in range(len(op.angles)) if (i, j) in param_map]
                if idx:
                    params = [p for _, p in idx]
                    dmats = derivative_matrices(op, ins.target, [j for j, _ in idx])
                    ol._gen.append((ol.n, ins, dict(idx)))
            ol.add_call(call, params, dmats)
        elif isinstance(op, Superoperator):
            if kind != L.BSIM_DM:
                raise TypeError("superoperators need a DensityMatrixSimulator")
            ol.add(L.OP_SUPEROP, list(reversed(ins.target)), op.matrix)  # bsim: matrix bit b <-> targets[b]
        elif isinstance(op, Noise):
            if kind != L.BSIM_DM:
                raise TypeError("noise channels need a DensityMatrixSimulator")
            if isinstance(op, PhaseFlip):
                ol.add(L.OP_PHASEFLIP, ins.target, np.array([op.probability], dtype=np.complex128))
            else:
                ks = op.kraus_rep()
                flat = np.concatenate([L.cmat(k).ravel(order="F") for k in ks])
                typ = L.OP_KRAUS_PHYS if noise_mode == "physical" else L.OP_KRAUS_REF
                ol.add(typ, ins.target, flat, n_kraus=len(ks))
        else:
            raise TypeError(f"unsupported operator {op!r}")
    return ol


def _kron_to_ascending(m: np.ndarray, targets) -> np.ndarray:
    """Re-index a matrix in kron order of ``targets`` (first = MSB) to kron
    order of the ascending targets."""
    k = len(targets)
    order = sorted(range(k), key=lambda i: targets[i])
    t = m.reshape([2] * (2 * k))
    perm = order + [k + o for o in order]
    return t.transpose(perm).reshape(2 ** k, 2 ** k)


def _reorder_kron(m: np.ndarray, src, dst) -> np.ndarray:
    """Re-index a matrix in kron order of qubits ``src`` (first = MSB) to
    kron order of ``dst`` (the same qubits in another order)."""
    k = len(src)
    perm = [list(src).index(q) for q in dst]
    t = m.reshape([2] * (2 * k))
    return t.transpose(perm + [k + p for p in perm]).reshape(2 ** k, 2 ** k)


def _jl_shl(x: int, s: int) -> int:
    """Julia's x << s (a negative s shifts right)."""
    return x << s if s >= 0 else x >> -s


def _swap_bits(ix: int, qubit_map: dict) -> int:
    """swap_bits (src/dm_simulator.jl:264-278), Julia shift semantics kept."""
    for in_q, out_q in qubit_map.items():
        if in_q < out_q:
            in_val = _jl_shl(_jl_shl(1, in_q) & ix, -in_q)
            out_val = _jl_shl(_jl_shl(1, out_q) & ix, -out_q)
            if in_val != out_val:
                ix = ix ^ _jl_shl(1, in_q) ^ _jl_shl(1, out_q)
    return ix


def _seed() -> int:
    return int.from_bytes(os.urandom(8), "little")


class _Simulator:
    kind = L.BSIM_SV

    def __init__(self, qubit_count: int, shots: int = 0):
        self._h = C.c_void_p()
        self.qubit_count = int(qubit_count)
        self.shots = int(shots)
        # 0 qubits (the reference's create_sim, src/BraketSimulator.jl:402-408):
        # no device state until reinit gives the program's qubit count
        if self.qubit_count != 0:
            L.call("bsim_state_create", self.qubit_count, self.kind, C.byref(self._h))
        self._gen = 0  # bumped whenever the device state is re-created
        self._observed = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.load().bsim_state_destroy(h)
            except Exception:
                pass
            self._h = None
        obs = getattr(self, "_observed", None)
        if obs is not None:
            obs.__del__()

    # -- lifecycle ----------------------------------------------------------
    def reinit(self, qubit_count: int, shots: int):
        """reinit! (src/sv_simulator.jl:96-127, src/dm_simulator.jl:96-128)."""
        if qubit_count != self.qubit_count or not self._h.value:
            if self._h.value:
                L.load().bsim_state_destroy(self._h)
            self._h = C.c_void_p()
            self._gen += 1
            self.qubit_count = int(qubit_count)
            L.call("bsim_state_create", self.qubit_count, self.kind, C.byref(self._h))
        else:
            L.call("bsim_state_reset", self._h)
        self.shots = int(shots)
        self._observed = None
        return self

    def similar(self, shots=None):
        return type(self)(self.qubit_count, self.shots if shots is None else shots)

    @property
    def handle(self):
        return self._h

    def _len(self) -> int:
        return 2 ** (self.qubit_count if self.kind == L.BSIM_SV else 2 * self.qubit_count)

    def upload(self, data: np.ndarray):
        a = np.ascontiguousarray(np.asarray(data, dtype=np.complex128).ravel(order="F"))
        L.call("bsim_state_upload", self._h, a.ctypes.data_as(C.POINTER(C.c_double)), a.size)

    def _download(self) -> np.ndarray:
        out = np.empty(self._len(), dtype=np.complex128)
        L.call("bsim_state_download", self._h, out.ctypes.data_as(C.POINTER(C.c_double)), out.size)
        return out

    # -- evolution -------------------------------------------------------------
    def evolve(self, operations, mode: int = 0):
        """evolve!: apply ``operations`` (Instructions) in order.  mode 0 fuses
        gates into LDS-tiled HBM passes, mode 1 launches one kernel per op."""
        ol = lower_instructions(list(operations), self.kind, getattr(self, "noise_mode", "reference"))
        L.call("bsim_evolve", self._h, ol.ptr, ol.n, int(mode))
        return self

    def partial_trace(self, output_qubits=None):
        """partial_trace(sim, output_qubits) (src/sv_simulator.jl:272-316,
        src/dm_simulator.jl:280-323), with the reference's special cases:
        no outputs -> the norm (SV, real) or trace (DM); every qubit -> the
        full matrix, unpermuted; unsorted outputs -> the sorted reduction
        reindexed by the reference's swap_bits rule.  The reduction itself
        runs on the device (bsim_partial_trace)."""
        n = self.qubit_count
        out_q = list(range(n)) if output_qubits is None else [int(q) for q in output_qubits]
        if not out_q:
            v = self._reduced([])[0, 0]
            return float(v.real) if self.kind == L.BSIM_SV else complex(v)
        if len(set(out_q)) == n:
            return self.density_matrix()
        k = len(out_q)
        red = self._reduced(sorted(out_q))
        if out_q == sorted(out_q):
            return red
        mapping = {k - q - 1: k - m - 1 for m, q in enumerate(out_q)}  # output_qubit_mapping
        D = 2 ** k
        out = np.zeros((D, D), dtype=np.complex128)
        sw = [_swap_bits(i, mapping) for i in range(D)]
        if self.kind == L.BSIM_SV:  # final_rho[out_ix, out_jx] += traced over ix <= jx
            for ix in range(D):
                for jx in range(ix, D):
                    out[sw[ix], sw[jx]] += red[ix, jx]
                    if sw[jx] != sw[ix]:
                        out[sw[jx], sw[ix]] += np.conj(red[ix, jx])
        else:  # final_rho[out_ix, out_jx] = sum(...) in raw_ix order
            for raw in range(D * D):
                ix, jx = raw // D, raw % D
                out[sw[ix], sw[jx]] = red[ix, jx]
        return out

    def _reduced(self, targets):
        k = len(targets)
        out = np.empty((2 ** k) * (2 ** k), dtype=np.complex128)
        t = np.array(targets, dtype=np.int32)
        L.call("bsim_partial_trace", self._h, k, L.iptr(t), L.dptr(out))
        return out.reshape((2 ** k, 2 ** k), order="F")

    def capture(self, operations, reset: bool = True) -> "CapturedEvolution":
        """Record evolve!(operations) -- preceded by the reset to |0...0>
        when ``reset`` -- as one hipGraph on this state (bsim_graph_create)."""
        return CapturedEvolution(self, operations, reset)

    def autotune(self, operations) -> dict:
        """Time the fused evolution of ``operations`` for every tile
        configuration, then every fusion width at the best one, and keep the
        fastest (bsim_autotune).  Returns {tile_bits, low_bits, sweep_bits,
        fusion_limit, merge_width, ms}; the state is left unspecified."""
        ol = lower_instructions(list(operations), self.kind, getattr(self, "noise_mode", "reference"))
        t, lo, rb, fl, ms = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_double()
        L.call("bsim_autotune", self._h, ol.ptr, ol.n, C.byref(t), C.byref(lo), C.byref(rb), C.byref(fl),
               C.byref(ms))
        mw = C.c_int()
        L.call("bsim_set_merge_width", -1, C.byref(mw))
        return {"tile_bits": t.value, "low_bits": lo.value, "sweep_bits": rb.value, "fusion_limit": fl.value,
                "merge_width": mw.value, "ms": ms.value}

    def _apply_call(self, call: KernelCall, vq_targets, conj: int):
        lib = L.load()
        t = np.array(vq_targets, dtype=np.int32)
        k = len(t)
        if call.double_excitation is not None:
            orig = np.array(list(reversed(list(vq_targets))), dtype=np.int32)
            L.check(lib.bsim_apply_double_excitation(self._h, L.iptr(orig), call.double_excitation), "double_excitation")
            return
        if call.diagonal:
            d = np.ascontiguousarray(call.matrix, dtype=np.complex128)
            L.check(lib.bsim_apply_diagonal(self._h, k, L.iptr(t), L.dptr(d), conj), "apply_diagonal")
            return
        m = L.cmat(call.matrix)
        if call.controls:
            nq_shift = vq_targets[0] - call.targets[0]
            c = np.array([x + nq_shift for x in call.controls], dtype=np.int32)
            cv = np.array(call.cvals, dtype=np.int32)
            L.check(lib.bsim_apply_controlled(self._h, len(c), L.iptr(c), L.iptr(cv), k, L.iptr(t), L.dptr(m), conj),
                    "apply_controlled")
        elif k == 1:
            L.check(lib.bsim_apply_1q(self._h, int(t[0]), L.dptr(m), conj), "apply_1q")
        elif k == 2:
            L.check(lib.bsim_apply_2q(self._h, int(t[0]), int(t[1]), L.dptr(m), conj), "apply_2q")
        else:
            L.check(lib.bsim_apply_kq(self._h, k, L.iptr(t), L.dptr(m), conj), "apply_kq")

    # -- results -----------------------------------------------------------------
    def probabilities(self) -> np.ndarray:
        out = np.empty(2 ** self.qubit_count, dtype=np.float64)
        L.call("bsim_probabilities", self._h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def marginal_probabilities(self, targets) -> np.ndarray:
        t = np.array(targets, dtype=np.int32)
        out = np.empty(2 ** len(t), dtype=np.float64)
        L.call("bsim_marginal_probabilities", self._h, len(t), L.iptr(t), out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def expect_pauli_words(self, words) -> np.ndarray:
        words = list(words)
        out = np.empty(len(words), dtype=np.float64)
        if words:
            buf = "".join(words).encode()
            L.call("bsim_expect_pauli", self._h, len(words), buf, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def expect_z_all(self) -> np.ndarray:
        out = np.empty(self.qubit_count, dtype=np.float64)
        L.call("bsim_expect_z_all", self._h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def _expect_matrix(self, m: np.ndarray, targets) -> float:
        t = np.array(targets, dtype=np.int32)
        mm = L.cmat(m)
        out = C.c_double()
        L.call("bsim_expect_matrix", self._h, len(t), L.iptr(t), L.dptr(mm), C.byref(out))
        return out.value

    def _observable_matrix(self, obs: Observable, targets, squared=False):
        """Matrix and targets in the convention of the reference kernel that
        ``obs`` reaches (see expectation)."""
        targets = list(targets)
        if isinstance(obs, TensorProduct):
            fs = [f.matrix() for f in obs.factors]
            if squared:
                fs = [np.eye(f.shape[0], dtype=np.complex128) if o.pauli is not None or isinstance(o, HObs) else f @ f
                      for f, o in zip(fs, obs.factors)]
            # The reference applies each factor on its own targets
            # (src/result_types.jl:203-220); a Hermitian factor's matrix is in
            # kron order of its targets as given only for 2 qubits on a state
            # vector (src/sv_simulator.jl:179-198), otherwise of its targets
            # sorted (:200-230, src/dm_simulator.jl:190-224).  Bring every
            # factor to the kron order of its given targets first.
            ix = 0
            for j, (f, o) in enumerate(zip(fs, obs.factors)):
                kf = o.n_qubits
                tf = targets[ix:ix + kf]
                ix += kf
                if isinstance(o, HermitianObservable) and (kf >= 3 or (kf == 2 and self.kind != L.BSIM_SV)):
                    fs[j] = _reorder_kron(f, sorted(tf), tf)
            m = np.ones((1, 1), dtype=np.complex128)
            for f in fs:
                m = np.kron(m, f)
            sv_as_is = self.kind == L.BSIM_SV and len(targets) <= 2
            if not sv_as_is:
                m = _kron_to_ascending(m, targets)
            return m, targets
        m = obs.matrix()
        if squared:
            m = m @ m
        return m, targets

    def expectation(self, observable: Observable, *targets) -> float:
        """<O> on ``targets`` (src/sv_simulator.jl:240-248,
        src/dm_simulator.jl:253-260)."""
        targets = list(targets) if targets else list(range(self.qubit_count))
        if is_pauli_word(observable):
            if observable.pauli is not None and len(targets) != 1:
                raise ValueError("a one-qubit observable takes one target")
            return float(self.expect_pauli_words([pauli_string(observable, targets, self.qubit_count)])[0])
        m, t = self._observable_matrix(observable, targets)
        return self._expect_matrix(m, t)

    def expectation_op_squared(self, observable: Observable, *targets) -> float:
        """src/result_types.jl:168-194."""
        if is_pauli_word(observable) or isinstance(observable, HObs):
            return 1.0
        m, t = self._observable_matrix(observable, list(targets), squared=True)
        return self._expect_matrix(m, t)

    def samples(self, seed: int | None = None, alias_mode: int = L.ALIAS_PARALLEL, wsum: float = 0.0,
                shots: int | None = None) -> np.ndarray:
        """Alias-method shots (src/result_types.jl:56-80), drawn on the GPU from
        Philox-4x32-10 uniforms keyed by ``seed``."""
        n = self.shots if shots is None else int(shots)
        L.call("bsim_alias_build", self._h, int(alias_mode), float(wsum))
        out = np.empty(n, dtype=np.int64)
        L.call("bsim_sample", self._h, n, C.c_uint64(_seed() if seed is None else int(seed) & (2 ** 64 - 1)), L.i64ptr(out))
        return out

    def alias_table(self):
        n = 2 ** self.qubit_count
        ap = np.empty(n, dtype=np.float64)
        al = np.empty(n, dtype=np.int32)
        L.call("bsim_alias_download", self._h, ap.ctypes.data_as(C.POINTER(C.c_double)),
               al.ctypes.data_as(C.POINTER(C.c_int32)))
        return ap, al

    def sample_from_draws(self, index: np.ndarray, u: np.ndarray) -> np.ndarray:
        index = np.ascontiguousarray(index, dtype=np.int64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.empty(index.size, dtype=np.int64)
        L.call("bsim_sample_from_draws", self._h, index.size, L.i64ptr(index),
               u.ctypes.data_as(C.POINTER(C.c_double)), L.i64ptr(out))
        return out

    # -- observables ------------------------------------------------------------
    def apply_observables(self, observables):
        """apply_observables! (src/sv_simulator.jl:256-265): diagonalizing gates
        on a copy of the state."""
        if self._observed is not None:
            raise RuntimeError("observables have already been applied.")
        ops = []
        for obs, targets in observables:
            ops.extend(diagonalizing_gates(obs, list(targets)))
        cp = type(self)(self.qubit_count, self.shots)
        L.call("bsim_state_copy", cp._h, self._h)
        cp.evolve(ops)
        self._observed = cp
        return self

    def state_with_observables(self):
        if self._observed is None:
            raise RuntimeError("observables have not been applied.")
        return self._observed._result_state()


def diagonalizing_gates(obs: Observable, targets):
    """src/observables.jl:220-241."""
    if isinstance(obs, IObs) or isinstance(obs, ZObs):
        return []
    if isinstance(obs, HObs):
        return [Instruction(Ry(-np.pi / 4.0), t) for t in targets]
    if isinstance(obs, XObs):
        return [Instruction(HGate(), t) for t in targets]
    if isinstance(obs, YObs):
        return [Instruction(Unitary(1 / np.sqrt(2) * np.array([[1.0, -1j], [1.0, 1j]])), t) for t in targets]
    if isinstance(obs, HermitianObservable):
        vecs = np.linalg.eigh(obs.mat)[1]
        if obs.mat.shape[0] == 2 ** len(targets):
            return [Instruction(Unitary(vecs), targets)]
        if obs.mat.shape[0] == 2 and len(targets) > 1:
            return [Instruction(Unitary(vecs), t) for t in targets]
    if isinstance(obs, TensorProduct):
        out = []
        for f, t in zip(obs.factors, targets):
            out.extend(diagonalizing_gates(f, [t] if np.ndim(t) == 0 else list(t)))
        return out
    raise TypeError(f"no diagonalizing gates for {obs!r}")


class StateVectorSimulator(_Simulator):
    """StateVectorSimulator{ComplexF64, <device vector>} (src/sv_simulator.jl:6-42)."""

    kind = L.BSIM_SV
    device_id = "braket_sv_v2"

    def state_vector(self) -> np.ndarray:
        return self._download()

    _result_state = state_vector

    def apply_gate(self, gate: Gate, *targets):
        """apply_gate!(gate, state_vector, targets...) -- one kernel launch."""
        call = kernel_call(gate, targets)
        self._apply_call(call, call.targets, 0)
        return self

    def density_matrix(self) -> np.ndarray:
        v = self.state_vector()
        return np.outer(v, np.conj(v))

    def __repr__(self):
        return f"StateVectorSimulator(qubit_count={self.qubit_count}, shots={self.shots})"


class DensityMatrixSimulator(_Simulator):
    """DensityMatrixSimulator{ComplexF64, <device matrix>} (src/dm_simulator.jl:6-40).

    ``noise_mode`` "reference" reproduces apply_noise!(::Kraus) exactly
    (only row/column blocks that agree on the non-target bits are updated,
    src/noise_kernels.jl:105-196); "physical" applies the channel to every
    block."""

    kind = L.BSIM_DM
    device_id = "braket_dm_v2"

    def __init__(self, qubit_count: int, shots: int = 0, noise_mode: str = "reference"):
        super().__init__(qubit_count, shots)
        if noise_mode not in ("reference", "physical"):
            raise ValueError("noise_mode is 'reference' or 'physical'")
        self.noise_mode = noise_mode

    def similar(self, shots=None):
        return type(self)(self.qubit_count, self.shots if shots is None else shots, self.noise_mode)

    def density_matrix(self) -> np.ndarray:
        """The stored (column-major) matrix, as the reference returns it."""
        n = 2 ** self.qubit_count
        return self._download().reshape((n, n), order="F")

    _result_state = density_matrix

    def apply_gate(self, gate: Gate, *targets):
        """_evolve_op!(dms, gate, targets...) (src/dm_simulator.jl:130-140):
        U on the column bits, then conj(U) on the row bits of vec(rho)."""
        call = kernel_call(gate, targets)
        n = self.qubit_count
        self._apply_call(call, list(call.targets), 0)
        call2 = KernelCall(list(call.targets), call.matrix, list(call.controls), list(call.cvals), call.diagonal,
                           call.double_excitation)
        call2.controls = [c + n for c in call.controls]
        call2.targets = [t + n for t in call.targets]
        self._apply_call(call2, call2.targets, 1)
        return self

    def apply_noise(self, noise: Noise, *targets):
        """apply_noise! (src/noise_kernels.jl:7-23, :102-196)."""
        t = np.array(targets, dtype=np.int32)
        if isinstance(noise, PhaseFlip):
            L.call("bsim_dm_phaseflip", self._h, int(t[0]), float(noise.probability))
            return self
        ks = noise.kraus_rep()
        flat = np.concatenate([L.cmat(k).ravel(order="F") for k in ks])
        L.call("bsim_dm_kraus", self._h, len(t), L.iptr(t), len(ks), L.dptr(flat),
               1 if self.noise_mode == "physical" else 0)
        return self

    def apply_superop(self, superop, *targets):
        """One superoperator pass on ``targets`` (bsim_dm_apply_superop)."""
        sop = superop if isinstance(superop, Superoperator) else Superoperator(superop)
        if len(targets) != sop.qubit_count:
            raise ValueError("superoperator arity and target count differ")
        t = np.array(list(reversed(targets)), dtype=np.int32)  # bsim: matrix bit b <-> targets[b]
        m = L.cmat(sop.matrix)
        L.call("bsim_dm_apply_superop", self._h, len(t), L.iptr(t), L.dptr(m))
        return self

    def __repr__(self):
        return f"DensityMatrixSimulator(qubit_count={self.qubit_count}, shots={self.shots})"


class CapturedEvolution:
    """An evolve! recorded once as a hipGraph and replayed with ``launch``
    (bsim_graph_create / bsim_graph_launch / bsim_graph_destroy).  Only
    circuits whose fused plan is made of tile passes are captured (gates on
    at most 2 qubits, diagonals on at most 4 vector bits and, on a density
    matrix, PhaseFlip and 1-qubit physical-mode noise); others raise
    BsimError.  The graph holds its simulator and is destroyed
    by ``close`` (or when collected)."""

    def __init__(self, sim: _Simulator, operations, reset: bool = True):
        self._g = C.c_void_p()
        self._sim = sim
        self._gen = sim._gen
        ol = lower_instructions(list(operations), sim.kind, getattr(sim, "noise_mode", "reference"))
        L.call("bsim_graph_create", sim.handle, ol.ptr, ol.n, int(bool(reset)), C.byref(self._g))

    def launch(self) -> _Simulator:
        if not self._g.value:
            raise L.BsimError("captured evolution is closed")
        if self._sim.handle is None or self._sim._gen != self._gen:
            raise L.BsimError("the simulator's state was re-created after the capture")
        # libbsim checks again: a graph whose state was destroyed or re-created
        # (even at the same address) is refused
        L.call("bsim_graph_launch", self._g)
        return self._sim

    def close(self):
        g = getattr(self, "_g", None)
        if g is not None and g.value:
            try:
                L.load().bsim_graph_destroy(g)
            except Exception:
                pass
            self._g = C.c_void_p()

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
