"""ctypes binding of libperegrine_hip (include/peregrine_hip.h).

The struct layouts here are the C ABI's; `tests/test_abi.py` checks every
declared symbol is exported and that the sizes agree with the header.
The library is loaded from this package directory only; a missing library
is an error, never a silent fallback.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libperegrine_hip.so"

PG_OK = 0
PG_ERR_ARG = -1
PG_ERR_HIP = -2
PG_ERR_UNSUPPORTED = -3
PG_ERR_NOMEM = -4
PG_ERR_STATE = -5

PG_KERNEL_AUTO = 0
PG_KERNEL_GLOBAL = 1
PG_KERNEL_LDS = 2
PG_KERNEL_FAST = 3

PG_VSTORE_F32 = 0
PG_VSTORE_F16_TILE = 1
PG_VSTORE_U16_TILE = 2
PG_F64_SWEEP_PLAIN = 0
PG_SWEEP_JACOBI = 0
PG_SWEEP_IN_PLACE = 1
PG_F64_SWEEP_DIAGONAL = 1


class PgGrid(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32 * 6), ("lo", ctypes.c_float * 6), ("step", ctypes.c_float * 6)]


class PgModel(ctypes.Structure):
    _fields_ = [
        ("B", ctypes.c_float * 9),
        ("A", ctypes.c_float * 9),
        ("q", ctypes.c_float * 3),
        ("S", ctypes.c_float * 9),
        ("D", ctypes.c_float * 9),
        ("c", ctypes.c_float * 3),
        ("eps", ctypes.c_float * 3),
        ("knee", ctypes.c_float),
    ]


class PgParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_float),
        ("nsub", ctypes.c_int32),
        ("checks", ctypes.c_int32),
        ("adaptive_checks", ctypes.c_int32),
        ("nearest", ctypes.c_int32),
        ("cap", ctypes.c_float),
        ("ntau", ctypes.c_int32),
        ("tau_ratio", ctypes.c_float),
        ("cfl", ctypes.c_float),
        ("tau_min", ctypes.c_float),
        ("tau_max", ctypes.c_float),
        ("hmax", ctypes.c_float),
        ("ncoarse", ctypes.c_int32),
        ("rounds", ctypes.c_int32),
        ("delta0", ctypes.c_float),
        ("rk2", ctypes.c_int32),
        ("vclamp", ctypes.c_int32),
        ("simplex", ctypes.c_int32),
    ]


# name -> (restype, argtypes); the exported C ABI, in header order
_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
SIGNATURES = {
    "pg_version": (ctypes.c_char_p, []),
    "pg_abi_version": (ctypes.c_int, []),
    "pg_last_error": (ctypes.c_char_p, []),
    "pg_device_count": (ctypes.c_int, [ctypes.POINTER(_I32)]),
    "pg_create": (ctypes.c_int, [_I32, ctypes.POINTER(_P)]),
    "pg_destroy": (ctypes.c_int, [_P]),
    "pg_device_budget": (ctypes.c_int, [_P, ctypes.c_double, ctypes.POINTER(_I64)]),
    "pg_grid_fits": (ctypes.c_int, [_P, ctypes.POINTER(PgGrid), _I32, ctypes.c_double,
                                    ctypes.POINTER(_I32), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]),
    "pg_solver_create": (ctypes.c_int, [_P, ctypes.POINTER(PgGrid), ctypes.POINTER(PgModel),
                                        ctypes.POINTER(PgParams), _P, _I32, _P, _I32,
                                        ctypes.POINTER(_P)]),
    "pg_solver_destroy": (ctypes.c_int, [_P]),
    "pg_fill_seed": (ctypes.c_int, [_P, _P, _I64]),
    "pg_sweep": (ctypes.c_int, [_P, _I32, _I32, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_float)]),
    "pg_solve": (ctypes.c_int, [_P, _I32, ctypes.c_double, ctypes.c_float, ctypes.POINTER(_I32),
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)]),
    "pg_polish": (ctypes.c_int, [_P, _I32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)]),
    "pg_policy": (ctypes.c_int, [_P, _P]),
    "pg_download_V": (ctypes.c_int, [_P, _P, _I64, _I64]),
    "pg_download_V16": (ctypes.c_int, [_P, _P, _I64, _I64]),
    "pg_solver_create_ooc": (ctypes.c_int, [_P, ctypes.POINTER(PgGrid), ctypes.POINTER(PgModel),
                                            ctypes.POINTER(PgParams), _P, _I32, _P, _I32, _I64, _P,
                                            ctypes.POINTER(_P)]),
    "pg_ooc_tile": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P]),
    "pg_ooc_info": (ctypes.c_int, [_P, _P, ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                                   ctypes.POINTER(ctypes.c_double)]),
    "pg_solve_f64": (ctypes.c_int, [_P, _I32, _I32, ctypes.c_double, ctypes.POINTER(_I32),
                                    ctypes.POINTER(ctypes.c_double)]),
    "pg_sweep_f64": (ctypes.c_int, [_P, _I32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "pg_residual_f64": (ctypes.c_int, [_P, _I32, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]),
    "pg_download_V64": (ctypes.c_int, [_P, _P, _I64, _I64]),
    "pg_policy_f64": (ctypes.c_int, [_P, _P]),
    "pg_release_f64": (ctypes.c_int, [_P]),
    "pg_adjoint": (ctypes.c_int, [_P, _P, _I64, _I32, ctypes.c_double, _P, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(_I32)]),
    "pg_upload_V": (ctypes.c_int, [_P, _P, _I64, _I64]),
    "pg_solver_set_kernel": (ctypes.c_int, [_P, _I32]),
    "pg_solver_set_sweep_mode": (ctypes.c_int, [_P, _I32]),
    "pg_solver_kernel_info": (ctypes.c_int, [_P, ctypes.POINTER(_I32), ctypes.POINTER(ctypes.c_double)]),
    "pg_solver_probe_info": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "pg_last_sweep_ms": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float)]),
    "pg_solver_set_active": (ctypes.c_int, [_P, _I32]),
    "pg_escape_sweep": (ctypes.c_int, [_P, _I32, _I32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)]),
    "pg_solve_escape": (ctypes.c_int, [_P, _I32, ctypes.c_double, ctypes.c_float, ctypes.POINTER(_I32),
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)]),
    "pg_last_sweep_boxes": (ctypes.c_int, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "pg_solver_set_warm_start": (ctypes.c_int, [_P, _I32]),
    "pg_download_warm_start": (ctypes.c_int, [_P, _P]),
    "pg_solver_ncells": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
    "pg_sweep_noalloc": (ctypes.c_int, [_P, _P, _P, _P, _P, _I32, _P, _P, _P, _I32, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_float)]),
}

# pg_sweep_noalloc's static scratch per device, and the largest grid it takes
PG_NOALLOC_SCRATCH_BYTES = 32 << 20
PG_NOALLOC_MAX_CELLS = ((PG_NOALLOC_SCRATCH_BYTES - 16384) // 8) * 256


class PeregrineError(RuntimeError):
    """A libperegrine_hip call returned a non-zero status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


class Unsupported(PeregrineError):
    """The parameter combination is not built into the library."""


_lib = None


def _share_torch_runtime() -> None:
    """Let torch (when it is installed) bring up its HIP runtime first.

    torch ships its own libamdhip64 with the same soname as the system one.
    Loaded first, this library binds the system runtime and torch then loads
    a second runtime that sees no GPU; loaded second, it binds torch's, and
    both share one runtime. Without torch nothing changes.
    """
    try:
        import torch
    except Exception:
        return
    try:
        if getattr(torch.version, "hip", None):
            torch.cuda.is_available()
    except Exception:
        pass


def load(path: os.PathLike | str | None = None) -> ctypes.CDLL:
    """Load libperegrine_hip.so from the package directory (or `path`)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(
            f"{p} is not built; run __graft_entry__.build() (make -C csrc) first")
    _share_torch_runtime()
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(lib: ctypes.CDLL, fn: str, rc: int) -> None:
    if rc != PG_OK:
        msg = lib.pg_last_error().decode(errors="replace")
        cls = Unsupported if rc == PG_ERR_UNSUPPORTED else PeregrineError
        raise cls(fn, rc, msg)
