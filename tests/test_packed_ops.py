"""The compact op-list image (bsim_ops_pack / bsim_ops_unpack, host code):
round trips of state-vector lists with controls and derivative matrices
and of density-matrix lists with Kraus channels, PhaseFlip and a
superoperator, field by field and value by value; malformed images are
refused with BSIM_EINVAL.  No GPU is needed."""
import ctypes as C
import struct

import numpy as np
import pytest

from bsim import _lib as L
from bsim import gates as G
from bsim import noises as Nz
from bsim import workloads as W
from bsim.simulators import Instruction, OpImage, Superoperator, lower_instructions

from circuits import random_circuit

HDR = 24
REC = 160


def _values(op, count):
    return np.ctypeslib.as_array(op.matrix, shape=(2 * count,)).copy() if count else np.zeros(0)


def _counts(op):
    k = op.n_targets
    return {L.OP_GATE: 4 ** k, L.OP_DIAG: 2 ** k, L.OP_KRAUS_REF: op.n_kraus * 4 ** k,
            L.OP_KRAUS_PHYS: op.n_kraus * 4 ** k, L.OP_PHASEFLIP: 1, L.OP_SUPEROP: 16 ** k}[op.type]


def _same(ol, im):
    a = ol.ptr
    b = im.ptr
    assert im.n == ol.n
    for i in range(ol.n):
        x, y = a[i], b[i]
        for f in ("type", "n_targets", "n_controls", "n_kraus", "n_params"):
            assert getattr(x, f) == getattr(y, f), (i, f)
        assert list(x.targets)[:x.n_targets] == list(y.targets)[:y.n_targets]
        assert list(x.controls)[:x.n_controls] == list(y.controls)[:y.n_controls]
        assert list(x.control_values)[:x.n_controls] == list(y.control_values)[:y.n_controls]
        n = _counts(x)
        if x.type == L.OP_PHASEFLIP:
            assert x.matrix[0] == y.matrix[0] and y.matrix[1] == 0.0
        else:
            assert _values(x, n).tobytes() == _values(y, n).tobytes(), i
        for j in range(x.n_params):
            assert x.param_index[j] == y.param_index[j]
            d = 4 ** x.n_targets
            assert (np.ctypeslib.as_array(x.dmatrix[j], shape=(2 * d,)).tobytes()
                    == np.ctypeslib.as_array(y.dmatrix[j], shape=(2 * d,)).tobytes())


def test_round_trip_state_vector():
    rng = np.random.default_rng(1)
    ins = random_circuit(7, 120, rng)
    ol = lower_instructions(ins, L.BSIM_SV)
    _same(ol, OpImage.pack(ol))
    hins, pmap, _ = W.hea(8, 3, seed=2)
    ol = lower_instructions(hins, L.BSIM_SV, param_map=pmap)
    im = OpImage.pack(ol)
    _same(ol, im)
    # the image survives a trip through plain bytes
    _same(ol, OpImage(im.tobytes()))


def test_round_trip_density_matrix():
    rng = np.random.default_rng(3)
    n = 4
    ins = random_circuit(n, 20, rng)
    ins += [Instruction(Nz.Depolarizing(0.1), [1]), Instruction(Nz.TwoQubitDepolarizing(0.05), [0, 2]),
            Instruction(Nz.PhaseFlip(0.3), [3]), Instruction(Nz.AmplitudeDamping(0.2), [2])]
    u = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))[0]
    ins.append(Instruction(Superoperator(np.kron(u.conj(), u)), [1]))
    for mode in ("reference", "physical"):
        ol = lower_instructions(ins, L.BSIM_DM, mode)
        _same(ol, OpImage.pack(ol))


def _image():
    hins, pmap, _ = W.hea(5, 1, seed=4)
    return bytearray(OpImage.pack(lower_instructions(hins, L.BSIM_SV, param_map=pmap)).tobytes())


def _refused(buf, what):
    with pytest.raises(L.BsimError, match="unpack"):
        OpImage(bytes(buf))


def test_malformed_images_are_refused():
    good = _image()
    OpImage(bytes(good))
    n_ops = struct.unpack_from("<I", good, 8)[0]
    n_values = struct.unpack_from("<Q", good, 16)[0]
    assert len(good) == HDR + REC * n_ops + 16 * n_values

    b = bytearray(good); b[0:4] = b"XXXX"; _refused(b, "magic")
    b = bytearray(good); struct.pack_into("<I", b, 4, 2); _refused(b, "version")
    b = bytearray(good); struct.pack_into("<I", b, 12, REC + 8); _refused(b, "record size")
    _refused(good[:-8], "truncated values")
    _refused(good[:HDR + REC - 1], "truncated records")
    _refused(good[:10], "truncated header")
    b = bytearray(good); struct.pack_into("<I", b, 8, 1 << 30); _refused(b, "op count")
    b = bytearray(good); struct.pack_into("<Q", b, 16, 1 << 60); _refused(b, "value count")
    # op 0: matrix index past the values, a negative index, a derivative past the end
    b = bytearray(good); struct.pack_into("<q", b, HDR + 112, n_values); _refused(b, "matrix")
    b = bytearray(good); struct.pack_into("<q", b, HDR + 112, -1); _refused(b, "matrix")
    first_param = next(i for i in range(n_ops) if struct.unpack_from("<i", good, HDR + REC * i + 120)[0] > 0)
    b = bytearray(good); struct.pack_into("<q", b, HDR + REC * first_param + 136, n_values - 1); _refused(b, "dmatrix")
    # counts out of range: targets, controls, type, parameters
    for off, val in ((4, 0), (4, 9), (8, -1), (8, 9), (0, 9), (120, 4)):
        b = bytearray(good); struct.pack_into("<i", b, HDR + off, val); _refused(b, f"field {off}")
    # a Kraus op without matrices
    ol = lower_instructions([Instruction(Nz.BitFlip(0.1), [0])], L.BSIM_DM)
    b = bytearray(OpImage.pack(ol).tobytes()); struct.pack_into("<i", b, HDR + 12, 0); _refused(b, "n_kraus")


def test_pack_refuses_malformed_ops():
    ops = (L.Op * 1)()
    ops[0].type = L.OP_GATE
    ops[0].n_targets = 1
    size = C.c_size_t()
    with pytest.raises(L.BsimError, match="null matrix"):
        L.call("bsim_ops_pack", ops, 1, None, 0, C.byref(size))
    m = np.eye(2, dtype=np.complex128)
    ops[0].matrix = L.dptr(m)
    L.call("bsim_ops_pack", ops, 1, None, 0, C.byref(size))
    assert size.value == HDR + REC + 16 * 4
    buf = np.zeros(size.value, dtype=np.uint8)
    with pytest.raises(L.BsimError, match="too small"):
        L.call("bsim_ops_pack", ops, 1, buf.ctypes.data, size.value - 1, C.byref(size))
    ops[0].n_targets = 9
    with pytest.raises(L.BsimError, match="out of range"):
        L.call("bsim_ops_pack", ops, 1, None, 0, C.byref(size))
