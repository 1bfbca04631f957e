"""Host driver over libperegrine_hip, mirroring the reference's solve loop.

`HipSolver` owns one resident grid on the device (the CuArray uploads of
solver/src/Run.jl:1776-1784). `solve_value` keeps `solve_value!`'s meaning
(solver/src/Run.jl:1434-1474): fill V with cap, seed the target cells with
zero, sweep until quiet or out of iterations, return (iterations, last
delta). `sweep_gpu` is one sweep (solver/src/Gpu.jl:27-37). `gpu_budget` and
`gpu_fits` are solver/src/Gpu.jl:181-209. There is no CPU fallback: without
the built library every entry point raises.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import abi
from .model import Grid6, Model, Params, ctl_soa


class Context:
    """A bound device with its solver stream."""

    def __init__(self, device: int = 0):
        self.lib = abi.load()
        self._h = ctypes.c_void_p()
        abi.check(self.lib, "pg_create", self.lib.pg_create(int(device), ctypes.byref(self._h)))
        self.device = device

    def close(self):
        if self._h:
            self.lib.pg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def gpu_budget(headroom: float = 0.85, ctx: Context | None = None) -> int:
    """Device bytes this run may use (Gpu.jl:181-185)."""
    ctx = ctx or default_context()
    out = ctypes.c_int64()
    abi.check(ctx.lib, "pg_device_budget", ctx.lib.pg_device_budget(ctx._h, headroom, ctypes.byref(out)))
    return int(out.value)


def gpu_fits(g: Grid6, headroom: float = 0.85, vstore: int = abi.PG_VSTORE_F32,
             ctx: Context | None = None):
    """(fits, need GiB, free GiB) for a whole-grid solve (Gpu.jl:204-209)."""
    ctx = ctx or default_context()
    fits = ctypes.c_int32()
    need = ctypes.c_double()
    free = ctypes.c_double()
    gc = g.c()
    abi.check(ctx.lib, "pg_grid_fits",
              ctx.lib.pg_grid_fits(ctx._h, ctypes.byref(gc), vstore, headroom, ctypes.byref(fits),
                                   ctypes.byref(need), ctypes.byref(free)))
    return bool(fits.value), need.value, free.value


@dataclass
class SweepResult:
    sum_decrease: float
    linf: float
    kernel_ms: float


class HipSolver:
    """One resident value table on the device."""

    def __init__(self, g: Grid6, m: Model, p: Params, ctl: np.ndarray, occ: np.ndarray,
                 vstore: int = abi.PG_VSTORE_F32, ctx: Context | None = None,
                 ooc_budget: int | None = None, ooc_tile=(0, 0, 0)):
        """A solver with its tables in HBM, or with ooc_budget (device bytes)
        an out-of-core one: tables in pinned host memory, swept by tiles of
        ooc_tile = (nx, ny, nheading) interior cells ((0, 0, 0): the largest
        that fits the budget)."""
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.g, self.m, self.p = g, m, p
        ctl = np.asarray(ctl, dtype=np.float32)
        if ctl.ndim == 2:
            soa = ctl_soa(ctl)
            nctl = ctl.shape[0]
        else:
            soa = np.ascontiguousarray(ctl)
            nctl = soa.size // 3
        occ = np.ascontiguousarray(np.asarray(occ, dtype=np.uint8).reshape(-1))
        if occ.size != g.n[0] * g.n[1] * g.n[2]:
            raise ValueError(f"occupancy has {occ.size} cells, grid needs {g.n[0] * g.n[1] * g.n[2]}")
        self.nctl = nctl
        self._gc, self._mc, self._pc = g.c(), m.c_struct(), p.c()
        self._soa, self._occ = soa, occ
        self._h = ctypes.c_void_p()
        if ooc_budget is None:
            abi.check(self.lib, "pg_solver_create",
                      self.lib.pg_solver_create(self.ctx._h, ctypes.byref(self._gc), ctypes.byref(self._mc),
                                                ctypes.byref(self._pc), soa.ctypes.data, nctl,
                                                occ.ctypes.data, vstore, ctypes.byref(self._h)))
        else:
            tile = np.ascontiguousarray(ooc_tile, dtype=np.int32)
            abi.check(self.lib, "pg_solver_create_ooc",
                      self.lib.pg_solver_create_ooc(self.ctx._h, ctypes.byref(self._gc), ctypes.byref(self._mc),
                                                    ctypes.byref(self._pc), soa.ctypes.data, nctl,
                                                    occ.ctypes.data, int(vstore), int(ooc_budget), tile.ctypes.data,
                                                    ctypes.byref(self._h)))
        self.ncells = g.ncells

    def close(self):
        if self._h:
            self.lib.pg_solver_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_kernel(self, kernel: int) -> None:
        """abi.PG_KERNEL_AUTO / _GLOBAL / _LDS (identical results)."""
        abi.check(self.lib, "pg_solver_set_kernel", self.lib.pg_solver_set_kernel(self._h, int(kernel)))

    def kernel_info(self):
        k = ctypes.c_int32()
        cov = ctypes.c_double()
        abi.check(self.lib, "pg_solver_kernel_info",
                  self.lib.pg_solver_kernel_info(self._h, ctypes.byref(k), ctypes.byref(cov)))
        return k.value, cov.value

    def probe_info(self):
        """(fraction of boxes whose probe chords meet no obstacle, fraction
        whose every pair passes its probes), from the exact plan."""
        f1 = ctypes.c_double()
        f2 = ctypes.c_double()
        abi.check(self.lib, "pg_solver_probe_info",
                  self.lib.pg_solver_probe_info(self._h, ctypes.byref(f1), ctypes.byref(f2)))
        return f1.value, f2.value

    # -- the reference's loop pieces ---------------------------------------
    def fill_seed(self, seeds) -> None:
        s = np.ascontiguousarray(np.asarray(seeds, dtype=np.int64))
        abi.check(self.lib, "pg_fill_seed", self.lib.pg_fill_seed(self._h, s.ctypes.data, s.size))

    def sweep(self, phase: int = 0, rev: bool = False) -> SweepResult:
        d = ctypes.c_double()
        l = ctypes.c_float()
        abi.check(self.lib, "pg_sweep",
                  self.lib.pg_sweep(self._h, int(phase), int(bool(rev)), ctypes.byref(d), ctypes.byref(l)))
        ms = ctypes.c_float()
        self.lib.pg_last_sweep_ms(self._h, ctypes.byref(ms))
        return SweepResult(d.value, l.value, ms.value)

    def solve(self, iters: int, tol: float = -1.0, tol_linf: float = 0.0):
        it = ctypes.c_int32()
        d = ctypes.c_double()
        l = ctypes.c_float()
        abi.check(self.lib, "pg_solve",
                  self.lib.pg_solve(self._h, int(iters), float(tol), float(tol_linf), ctypes.byref(it),
                                    ctypes.byref(d), ctypes.byref(l)))
        return it.value, d.value, l.value

    def escape_sweep(self, phase: int = 0, rev: bool = False) -> SweepResult:
        """One escape sweep (sweep_escape_gpu!): returns the summed and the
        largest improvement of the escape times."""
        d = ctypes.c_double()
        l = ctypes.c_float()
        abi.check(self.lib, "pg_escape_sweep",
                  self.lib.pg_escape_sweep(self._h, int(phase), int(bool(rev)), ctypes.byref(d), ctypes.byref(l)))
        ms = ctypes.c_float()
        self.lib.pg_last_sweep_ms(self._h, ctypes.byref(ms))
        return SweepResult(d.value, l.value, ms.value)

    def solve_escape(self, iters: int, tol: float = -1.0, tol_linf: float = 0.0):
        """solve_escape!: escape sweeps until quiet (sum <= tol, or max < tol_linf)."""
        it = ctypes.c_int32()
        d = ctypes.c_double()
        l = ctypes.c_float()
        abi.check(self.lib, "pg_solve_escape",
                  self.lib.pg_solve_escape(self._h, int(iters), float(tol), float(tol_linf), ctypes.byref(it),
                                           ctypes.byref(d), ctypes.byref(l)))
        return it.value, d.value, l.value

    def set_sweep_mode(self, mode: int) -> None:
        """abi.PG_SWEEP_JACOBI (default, deterministic) or abi.PG_SWEEP_IN_PLACE."""
        abi.check(self.lib, "pg_solver_set_sweep_mode", self.lib.pg_solver_set_sweep_mode(self._h, int(mode)))

    def set_warm_start(self, on: bool) -> None:
        """The Int8 warm-start policy store (Run.jl:1782, PeregrineSolver.jl:
        728-744): zeroed by fill_seed, read and written by every sweep."""
        abi.check(self.lib, "pg_solver_set_warm_start", self.lib.pg_solver_set_warm_start(self._h, int(bool(on))))

    def download_warm_start(self) -> np.ndarray:
        out = np.empty(3 * self.g.ncells, np.int8)
        abi.check(self.lib, "pg_download_warm_start", self.lib.pg_download_warm_start(self._h, out.ctypes.data))
        return out

    def set_active(self, on: bool) -> None:
        """Skip the boxes whose inputs did not change (default on; results
        are bit for bit those of computing every box)."""
        abi.check(self.lib, "pg_solver_set_active", self.lib.pg_solver_set_active(self._h, int(bool(on))))

    def last_sweep_boxes(self):
        """(boxes the last sweep computed, boxes of the grid)."""
        c = ctypes.c_int64()
        t = ctypes.c_int64()
        abi.check(self.lib, "pg_last_sweep_boxes",
                  self.lib.pg_last_sweep_boxes(self._h, ctypes.byref(c), ctypes.byref(t)))
        return c.value, t.value

    def polish(self, nsweeps: int):
        """Decode an FP16 store to FP32 and run `nsweeps` exact sweeps."""
        d = ctypes.c_double()
        l = ctypes.c_float()
        abi.check(self.lib, "pg_polish", self.lib.pg_polish(self._h, int(nsweeps), ctypes.byref(d), ctypes.byref(l)))
        return d.value, l.value

    def policy(self) -> np.ndarray:
        out = np.empty(self.ncells, dtype=np.uint8)
        abi.check(self.lib, "pg_policy", self.lib.pg_policy(self._h, out.ctypes.data))
        return out

    def download(self, off: int = 0, n: int | None = None) -> np.ndarray:
        n = self.ncells - off if n is None else n
        out = np.empty(n, dtype=np.float32)
        abi.check(self.lib, "pg_download_V", self.lib.pg_download_V(self._h, out.ctypes.data, off, n))
        return out

    def download_mirror(self, off: int = 0, n: int | None = None) -> np.ndarray:
        """The 16-bit mirror the precision form's gathers read, decoded."""
        n = self.ncells - off if n is None else n
        out = np.empty(n, dtype=np.float32)
        abi.check(self.lib, "pg_download_V16", self.lib.pg_download_V16(self._h, out.ctypes.data, off, n))
        return out

    def ooc_info(self):
        """(tile interior, tile count, device bytes, cells loaded per cell
        updated) of an out-of-core solver."""
        tile = np.zeros(3, np.int32)
        nt = ctypes.c_int64()
        nb = ctypes.c_int64()
        amp = ctypes.c_double()
        abi.check(self.lib, "pg_ooc_info", self.lib.pg_ooc_info(self._h, tile.ctypes.data, ctypes.byref(nt),
                                                                 ctypes.byref(nb), ctypes.byref(amp)))
        return tuple(int(v) for v in tile), nt.value, nb.value, amp.value

    def ooc_tiles(self) -> np.ndarray:
        """Per tile of an out-of-core solver: interior origin, interior
        extent, window origin, window extent (each x, y, heading), (n, 12)."""
        n = self.ooc_info()[1]
        out = np.zeros((n, 12), np.int32)
        for k in range(n):
            v = [np.zeros(3, np.int32) for _ in range(4)]
            abi.check(self.lib, "pg_ooc_tile", self.lib.pg_ooc_tile(self._h, k, *(a.ctypes.data for a in v)))
            out[k] = np.concatenate(v)
        return out

    # FP64 tables ---------------------------------------------------------
    def solve_f64(self, iters: int, tol_linf: float, diagonal: bool = True):
        """FP64 table from cap and the seeds; returns (sweeps, last ||dV||_inf)."""
        it = ctypes.c_int32()
        l = ctypes.c_double()
        mode = abi.PG_F64_SWEEP_DIAGONAL if diagonal else abi.PG_F64_SWEEP_PLAIN
        abi.check(self.lib, "pg_solve_f64", self.lib.pg_solve_f64(self._h, mode, int(iters), float(tol_linf),
                                                                   ctypes.byref(it), ctypes.byref(l)))
        return it.value, l.value

    def sweep_f64(self, diagonal: bool = True):
        """One more FP64 sweep; returns (||dV||_inf, summed decrease)."""
        l = ctypes.c_double()
        d = ctypes.c_double()
        mode = abi.PG_F64_SWEEP_DIAGONAL if diagonal else abi.PG_F64_SWEEP_PLAIN
        abi.check(self.lib, "pg_sweep_f64", self.lib.pg_sweep_f64(self._h, mode, ctypes.byref(l), ctypes.byref(d)))
        return l.value, d.value

    def residual_f64(self, which: int = 0):
        """(max, sum) of |T V - V| with T one FP64 backup; which 0 = FP64
        table, 1 = the current FP32 table."""
        l = ctypes.c_double()
        d = ctypes.c_double()
        abi.check(self.lib, "pg_residual_f64", self.lib.pg_residual_f64(self._h, int(which), ctypes.byref(l),
                                                                         ctypes.byref(d)))
        return l.value, d.value

    def download_f64(self, off: int = 0, n: int | None = None) -> np.ndarray:
        n = self.ncells - off if n is None else n
        out = np.empty(n, dtype=np.float64)
        abi.check(self.lib, "pg_download_V64", self.lib.pg_download_V64(self._h, out.ctypes.data, off, n))
        return out

    def policy_f64(self) -> np.ndarray:
        out = np.empty(self.ncells, dtype=np.uint8)
        abi.check(self.lib, "pg_policy_f64", self.lib.pg_policy_f64(self._h, out.ctypes.data))
        return out

    def adjoint(self, starts: np.ndarray, max_iters: int = 100000, tol: float = 1e-12):
        """(L, dL/dtheta (30), iterations) at the FP64 table; L is the mean
        multilinear V over the starts (rows of x, y, heading, vx, vy, omega)."""
        st = np.ascontiguousarray(starts, dtype=np.float64).reshape(-1, 6)
        grad = np.zeros(30, np.float64)
        L = ctypes.c_double()
        it = ctypes.c_int32()
        abi.check(self.lib, "pg_adjoint", self.lib.pg_adjoint(self._h, st.ctypes.data, st.shape[0], int(max_iters),
                                                               float(tol), grad.ctypes.data, ctypes.byref(L),
                                                               ctypes.byref(it)))
        return L.value, grad, it.value

    def release_f64(self) -> None:
        abi.check(self.lib, "pg_release_f64", self.lib.pg_release_f64(self._h))

    def upload(self, V: np.ndarray, off: int = 0) -> None:
        V = np.ascontiguousarray(np.asarray(V, dtype=np.float32))
        abi.check(self.lib, "pg_upload_V", self.lib.pg_upload_V(self._h, V.ctypes.data, off, V.size))


def sweep_gpu(solver: HipSolver, phase: int = 0, rev: bool = False) -> float:
    """One sweep; returns the total improvement (Gpu.jl:27-37)."""
    return solver.sweep(phase, rev).sum_decrease


def solve_value(solver: HipSolver, tidx, iters: int, tol: float, tol_linf: float = 0.0,
                on_progress=None):
    """solve_value! (Run.jl:1434-1474): returns (iterations_done, final_delta).

    `tidx` holds 0-based seed indices. Sweep i takes phase i - 1 (the
    rotating slice's offset) and alternates direction. The quiet rule is the
    reference's: stop after quiet_needed sweeps in a row whose summed
    decrease is <= tol, one for a full control scan and a whole turn of the
    slice otherwise (Run.jl:1440-1441); with tol_linf > 0 a sweep whose
    ||dV||_inf < tol_linf also counts as quiet. on_progress(it, result) is
    called on sweep 1, every 5th and the last.
    """
    p = solver.p
    need = 1 if p.ncoarse <= 0 else max(1, -(-(solver.nctl - 1) // p.ncoarse))
    solver.fill_seed(tidx)
    done, last, quiet = 0, float("inf"), 0
    for it in range(1, iters + 1):
        r = solver.sweep(it - 1, it % 2 == 1)
        done, last = it, r.sum_decrease
        q = (tol >= 0 and r.sum_decrease <= tol) or (tol_linf > 0 and r.linf < tol_linf)
        quiet = quiet + 1 if q else 0
        if on_progress is not None and (it % 5 == 0 or it == 1 or quiet >= need):
            on_progress(it, r)
        if quiet >= need:
            break
    return done, last
