"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, and only as the checker or the timed CPU baseline. The struct
types are the C ABI's (include/peregrine_hip.h), shared with the product so
one problem description feeds both.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_F = ctypes.c_float
_D = ctypes.c_double
SIG = {
    "orc_abi": (ctypes.c_int, []),
    "orc_sizeof": (_I64, [_I32]),
    "orc_sincosf": (None, [_F, _P, _P]),
    "orc_sincosf_batch": (None, [_P, _P, _P, _I64]),
    "orc_interp_f32": (_F, [_P, _P, _P, _I32, _F, _I32]),
    "orc_interp_f64": (_D, [_P, _P, _P, _I32, _D, _I32]),
    "orc_interp_kuhn_f32": (_F, [_P, _P, _P, _F, _I32, _P]),
    "orc_step_state_f32": (None, [_P, _P, _P, _F, _I32, _I32, _P]),
    "orc_step_state_f64": (None, [_P, _P, _P, _D, _I32, _I32, _P]),
    "orc_traction_gain": (_F, [_F, _F, _F, _F]),
    "orc_body_accel_f64": (None, [_P, _P, _P, _P]),
    "orc_body_accel_f32": (None, [_P, _P, _P, _P]),
    "orc_lookahead_f32": (_F, [_P, _P, _P, _P, _P, _P, _F, _I32, _P]),
    "orc_substeps_for": (_I32, [_F, _P]),
    "orc_cfl_tau": (_F, [_P, _P, _F, _F, _F, _P]),
    "orc_cell_update_f32": (_F, [_I64, _P, _P, _P, _P, _P, _I32, _P, _I32, _I32, _P]),
    "orc_cell_update_f64": (_D, [_I64, _P, _P, _P, _P, _P, _I32, _P, _I32, _I32, _P]),
    "orc_backup_cells_f64": (ctypes.c_int, [_P, _P, _P, _P, _P, _I32, _P, _P, _I64, _P, _P, _I32]),
    "orc_backup_cells_f32": (ctypes.c_int, [_P, _P, _P, _P, _P, _I32, _P, _P, _I64, _P, _I32]),
    "orc_sweep_f32": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _I32, _P, _I32, _I32, _I32, _I32, _P, _P, _P, _I32]),
    "orc_sweep_warm_f32": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _I32, _P, _I32, _I32, _I32, _I32, _P, _P, _P]),
    "orc_tanh_f32": (_F, [_F]),
    "orc_sweep_range_f32": (ctypes.c_int, [_P, _P, _P, _P, _P, _I32, _P, _I64, _I64, _I32, _I32, _I32, _P, _P]),
    "orc_sweep_f64": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _I32, _P, _I32, _I32, _I32, _I32, _P, _P, _P, _I32]),
    "orc_target_cells": (_I64, [_P, _P, _P, _P, _I64]),
    "orc_check_quotient": (_I64, [_F, _I32]),
    "orc_adjoint_f64": (_I32, [_P, _P, _P, _P, _P, _I32, _P, _P, _I64, _P, _I64, _I32, ctypes.c_double, _I32,
                               _P, _P, _P]),
}

_lib = None


def build() -> None:
    """Compile liboracle.so from the sources beside this file."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        for name, (res, args) in SIG.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


class Problem:
    """A problem description in C-ABI structs plus its arrays."""

    def __init__(self, gc, mc, pc, ctl_soa: np.ndarray, occ: np.ndarray):
        self.gc, self.mc, self.pc = gc, mc, pc
        self.ctl = np.ascontiguousarray(ctl_soa, dtype=np.float32)
        self.nctl = self.ctl.size // 3
        self.occ = np.ascontiguousarray(occ, dtype=np.uint8)
        self.ncell = int(np.prod([gc.n[k] for k in range(6)], dtype=np.int64))

    def sweep(self, Vin: np.ndarray, Vout: np.ndarray | None = None, mode: int = 0, rev: int = 0,
              phase: int = 0, nthreads: int = 0, pol: np.ndarray | None = None, esc: int = 0):
        """mode 0 Jacobi (Vout written), 1 serial in place, 2 threaded in place."""
        L = lib()
        d = ctypes.c_double()
        l = ctypes.c_double()
        f = L.orc_sweep_f64 if Vin.dtype == np.float64 else L.orc_sweep_f32
        if mode != 0:
            Vout = Vin
        rc = f(_ptr(Vin), _ptr(Vout), _ptr(self.occ), ctypes.byref(self.gc), ctypes.byref(self.mc),
               _ptr(self.ctl), self.nctl, ctypes.byref(self.pc), mode, rev, phase, nthreads,
               _ptr(pol), ctypes.byref(d), ctypes.byref(l), esc)
        if rc != 0:
            raise RuntimeError(f"oracle sweep failed ({rc})")
        return d.value, l.value

    def sweep_range(self, V: np.ndarray, lo: int, hi: int, rev: int = 0, phase: int = 0, nthreads: int = 0):
        """In place, threaded, over the cells [lo, hi) of float V only."""
        assert V.dtype == np.float32 and V.flags.c_contiguous and V.size == self.ncell
        d = ctypes.c_double()
        l = ctypes.c_double()
        rc = lib().orc_sweep_range_f32(_ptr(V), _ptr(self.occ), ctypes.byref(self.gc), ctypes.byref(self.mc),
                                       _ptr(self.ctl), self.nctl, ctypes.byref(self.pc), int(lo), int(hi),
                                       int(rev), int(phase), int(nthreads), ctypes.byref(d), ctypes.byref(l))
        if rc != 0:
            raise RuntimeError(f"oracle sweep failed ({rc})")
        return d.value, l.value

    def sweep_warm(self, Vin: np.ndarray, Vout: np.ndarray, wpol: np.ndarray, phase: int = 0, nthreads: int = 0):
        """A Jacobi float sweep reading and updating the warm-start store
        wpol (int8, 3 per cell, interleaved) in place."""
        assert wpol.dtype == np.int8 and wpol.size == 3 * self.ncell and wpol.flags.c_contiguous
        L = lib()
        d = ctypes.c_double()
        l = ctypes.c_double()
        rc = L.orc_sweep_warm_f32(_ptr(Vin), _ptr(Vout), _ptr(self.occ), ctypes.byref(self.gc), ctypes.byref(self.mc),
                                  _ptr(self.ctl), self.nctl, ctypes.byref(self.pc), 0, 0, phase, nthreads,
                                  _ptr(wpol), ctypes.byref(d), ctypes.byref(l))
        if rc != 0:
            raise RuntimeError(f"oracle sweep failed ({rc})")
        return d.value, l.value

    def cell_update(self, V: np.ndarray, idx: int, phase: int = 0, esc: int = 0):
        L = lib()
        k = ctypes.c_int32()
        f = L.orc_cell_update_f64 if V.dtype == np.float64 else L.orc_cell_update_f32
        v = f(int(idx), _ptr(V), _ptr(self.occ), ctypes.byref(self.gc), ctypes.byref(self.mc),
              _ptr(self.ctl), self.nctl, ctypes.byref(self.pc), phase, esc, ctypes.byref(k))
        return v, k.value

    def backup_cells_f64(self, V: np.ndarray, cells: np.ndarray, nthreads: int = 0):
        L = lib()
        cells = np.ascontiguousarray(cells, dtype=np.int64)
        out = np.empty(cells.size, dtype=np.float64)
        kout = np.empty(cells.size, dtype=np.int32)
        rc = L.orc_backup_cells_f64(_ptr(V), _ptr(self.occ), ctypes.byref(self.gc), ctypes.byref(self.mc),
                                    _ptr(self.ctl), self.nctl, ctypes.byref(self.pc), _ptr(cells),
                                    cells.size, _ptr(out), _ptr(kout), nthreads)
        if rc != 0:
            raise RuntimeError(f"oracle backup failed ({rc})")
        return out, kout

    def backup_cells_f32(self, V: np.ndarray, cells: np.ndarray, nthreads: int = 0):
        L = lib()
        cells = np.ascontiguousarray(cells, dtype=np.int64)
        out = np.empty(cells.size, dtype=np.float32)
        rc = L.orc_backup_cells_f32(_ptr(V), _ptr(self.occ), ctypes.byref(self.gc), ctypes.byref(self.mc),
                                    _ptr(self.ctl), self.nctl, ctypes.byref(self.pc), _ptr(cells),
                                    cells.size, _ptr(out), nthreads)
        if rc != 0:
            raise RuntimeError(f"oracle backup failed ({rc})")
        return out

    def adjoint_f64(self, V: np.ndarray, seeds, starts: np.ndarray, max_iters: int = 100000, tol: float = 1e-12,
                    nthreads: int = 0):
        """dL/dtheta (30) and L = mean interp(V, start) at the FP64 table V,
        with the argmin policy held fixed; returns (grad, L, iterations,
        policy). iterations is -1 if the adjoint did not converge."""
        L = lib()
        V = np.ascontiguousarray(V, dtype=np.float64)
        seeds = np.ascontiguousarray(seeds, dtype=np.int64)
        starts = np.ascontiguousarray(starts, dtype=np.float64).reshape(-1, 6)
        grad = np.zeros(30, np.float64)
        lv = np.zeros(1, np.float64)
        pol = np.zeros(self.ncell, np.int32)
        it = L.orc_adjoint_f64(_ptr(V), _ptr(self.occ), ctypes.byref(self.gc), ctypes.byref(self.mc),
                               _ptr(self.ctl), self.nctl, ctypes.byref(self.pc), _ptr(seeds), seeds.size,
                               _ptr(starts), starts.shape[0], int(max_iters), float(tol), int(nthreads),
                               _ptr(grad), _ptr(lv), _ptr(pol))
        return grad, float(lv[0]), int(it), pol

    def solve(self, seeds, iters: int, tol_linf: float = 1e-3, dtype=np.float32, nthreads: int = 0,
              mode: int = 0):
        """Jacobi (mode 0) or serial in-place (mode 1) solve from cap with
        zero seeds, stopping at sup-norm change < tol_linf."""
        cap = self.pc.cap
        V = np.full(self.ncell, cap, dtype=dtype)
        V[np.asarray(seeds, dtype=np.int64)] = 0
        W = np.empty_like(V) if mode == 0 else V
        it = 0
        linf = float("inf")
        for it in range(1, iters + 1):
            _, linf = self.sweep(V, W, mode=mode, rev=(it % 2 == 1), phase=it - 1, nthreads=nthreads)
            if mode == 0:
                V, W = W, V
            V[np.asarray(seeds, dtype=np.int64)] = 0
            if linf < tol_linf:
                break
        return V, it, linf


def sincosf(x: np.ndarray):
    x = np.ascontiguousarray(x, dtype=np.float32)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib().orc_sincosf_batch(x.ctypes.data, s.ctypes.data, c.ctypes.data, x.size)
    return s, c


def check_quotient(b: float, nthreads: int = 0) -> int:
    return int(lib().orc_check_quotient(float(np.float32(b)), nthreads))
