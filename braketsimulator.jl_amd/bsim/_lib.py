"""ctypes binding of libbsim (include/bsim.h).

The library is the only compute path: importing this module does not
require a GPU, but every simulator call goes through libbsim and raises
``BsimError`` when the library or the device is missing.  There is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = Path(os.environ.get("BSIM_LIB", PKG_ROOT / "lib" / "libbsim.so"))

BSIM_OK = 0
BSIM_EINVAL = -1
BSIM_ENOTSUP = -4
BSIM_SV = 0
BSIM_DM = 1

OP_GATE = 0
OP_DIAG = 1
OP_KRAUS_REF = 2
OP_KRAUS_PHYS = 3
OP_PHASEFLIP = 4
OP_SUPEROP = 5

ALIAS_PARALLEL = 0
ALIAS_REFERENCE = 1

MAX_TARGETS = 8
MAX_CONTROLS = 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)


class BsimError(RuntimeError):
    """A libbsim call failed (or the library is not available)."""


class Op(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("n_targets", C.c_int32),
        ("n_controls", C.c_int32),
        ("n_kraus", C.c_int32),
        ("targets", C.c_int32 * MAX_TARGETS),
        ("controls", C.c_int32 * MAX_CONTROLS),
        ("control_values", C.c_int32 * MAX_CONTROLS),
        ("matrix", _dp),
        ("n_params", C.c_int32),
        ("param_index", C.c_int32 * 3),
        ("dmatrix", _dp * 3),
    ]


class Task(C.Structure):
    _fields_ = [
        ("qubit_count", C.c_int32),
        ("n_ops", C.c_int32),
        ("ops", C.POINTER(Op)),
        ("shots", C.c_int64),
        ("seed", C.c_uint64),
        ("expect_z", _dp),
        ("samples", _i64p),
        ("n_words", C.c_int32),
        ("n_params", C.c_int32),
        ("paulis", C.c_char_p),
        ("coeffs", _dp),
        ("grad", _dp),
        ("expval", _dp),
    ]


# name -> (restype, argtypes)
_SIGS = {
    "bsim_version": (C.c_int, []),
    "bsim_last_error": (C.c_int, [C.c_char_p, C.c_int64]),
    "bsim_device_info": (C.c_int, [C.c_char_p, C.c_int64, _ip, _i64p]),
    "bsim_state_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "bsim_state_destroy": (C.c_int, [C.c_void_p]),
    "bsim_state_reset": (C.c_int, [C.c_void_p]),
    "bsim_state_qubits": (C.c_int, [C.c_void_p, _ip, _ip]),
    "bsim_state_upload": (C.c_int, [C.c_void_p, _dp, C.c_int64]),
    "bsim_state_download": (C.c_int, [C.c_void_p, _dp, C.c_int64]),
    "bsim_state_copy": (C.c_int, [C.c_void_p, C.c_void_p]),
    "bsim_synchronize": (C.c_int, [C.c_void_p]),
    "bsim_state_device_ptr": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "bsim_apply_1q": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_int]),
    "bsim_apply_2q": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, C.c_int]),
    "bsim_apply_controlled": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, C.c_int, _ip, _dp, C.c_int]),
    "bsim_apply_kq": (C.c_int, [C.c_void_p, C.c_int, _ip, _dp, C.c_int]),
    "bsim_apply_diagonal": (C.c_int, [C.c_void_p, C.c_int, _ip, _dp, C.c_int]),
    "bsim_apply_double_excitation": (C.c_int, [C.c_void_p, _ip, C.c_double]),
    "bsim_evolve": (C.c_int, [C.c_void_p, C.POINTER(Op), C.c_int, C.c_int]),
    "bsim_ops_pack": (C.c_int, [C.POINTER(Op), C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "bsim_ops_unpack": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(Op), C.c_int, _ip]),
    "bsim_evolve_packed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "bsim_plan_stats": (C.c_int, [C.c_int, C.c_int, C.POINTER(Op), C.c_int, _i64p]),
    "bsim_set_tile_config": (C.c_int, [C.c_int, C.c_int]),
    "bsim_get_tile_config": (C.c_int, [_ip, _ip]),
    "bsim_set_fusion_limit": (C.c_int, [C.c_int, _ip]),
    "bsim_set_profiling": (C.c_int, [C.c_int]),
    "bsim_kernel_stats": (C.c_int, [C.c_int, _i64p, _dp]),
    "bsim_stream_probe": (C.c_int, [C.c_int, C.c_int64, C.c_int, C.c_int, _dp]),
    "bsim_dm_kraus": (C.c_int, [C.c_void_p, C.c_int, _ip, C.c_int, _dp, C.c_int]),
    "bsim_dm_phaseflip": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "bsim_probabilities": (C.c_int, [C.c_void_p, _dp]),
    "bsim_marginal_probabilities": (C.c_int, [C.c_void_p, C.c_int, _ip, _dp]),
    "bsim_expect_pauli": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, _dp]),
    "bsim_expect_z_all": (C.c_int, [C.c_void_p, _dp]),
    "bsim_expect_matrix": (C.c_int, [C.c_void_p, C.c_int, _ip, _dp, _dp]),
    "bsim_expect_pauli_sum": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, _dp, _dp]),
    "bsim_alias_build": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "bsim_alias_download": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int32)]),
    "bsim_sample": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint64, _i64p]),
    "bsim_sample_from_draws": (C.c_int, [C.c_void_p, C.c_int64, _i64p, _dp, _i64p]),
    "bsim_shots_to_bits": (C.c_int, [C.c_void_p, C.c_int64, _i64p, C.c_int, _ip, C.POINTER(C.c_int8)]),
    "bsim_adjoint_gradient": (
        C.c_int,
        [C.c_void_p, C.POINTER(Op), C.c_int, C.c_int, C.c_char_p, _dp, C.c_int, _dp, _dp],
    ),
    "bsim_batch_run": (C.c_int, [C.POINTER(Task), C.c_int, _dp]),
    "bsim_set_checkpoint_interval": (C.c_int, [C.c_int, C.POINTER(C.c_int)]),
    "bsim_set_adjoint_fusion": (C.c_int, [C.c_int, C.POINTER(C.c_int)]),
    "bsim_dm_apply_superop": (C.c_int, [C.c_void_p, C.c_int, _ip, _dp]),
    "bsim_graph_create": (C.c_int, [C.c_void_p, C.POINTER(Op), C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "bsim_graph_launch": (C.c_int, [C.c_void_p]),
    "bsim_graph_destroy": (C.c_int, [C.c_void_p]),
    "bsim_autotune": (C.c_int, [C.c_void_p, C.POINTER(Op), C.c_int, _ip, _ip, _ip, _ip, _dp]),
    "bsim_set_max_blocks": (C.c_int, [C.c_int, _ip]),
    "bsim_set_merge_width": (C.c_int, [C.c_int, _ip]),
    "bsim_evolve_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Op), C.c_int]),
    "bsim_partial_trace": (C.c_int, [C.c_void_p, C.c_int, _ip, _dp]),
    "bsim_set_sweep_bits": (C.c_int, [C.c_int, _ip]),
}

_lib = None


def load(path: Path | str | None = None):
    """Load libbsim.so (once) and declare every entry point of bsim.h."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise BsimError(
            f"libbsim not found at {p}: build it with `python braketsimulator.jl_amd/build.py`"
        )
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def last_error() -> str:
    lib = load()
    buf = C.create_string_buffer(1024)
    lib.bsim_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int, what: str = "") -> None:
    if rc != BSIM_OK:
        raise BsimError(f"{what}: libbsim error {rc}: {last_error()}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


# ---- array helpers ----------------------------------------------------

def dptr(a: np.ndarray):
    assert a.dtype in (np.float64, np.complex128) and a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32
    return a.ctypes.data_as(_ip)


def i64ptr(a: np.ndarray):
    assert a.dtype == np.int64
    return a.ctypes.data_as(_i64p)


def cmat(m) -> np.ndarray:
    """Column-major complex128 copy of a square matrix (Julia SMatrix layout)."""
    return np.asfortranarray(np.asarray(m, dtype=np.complex128))
