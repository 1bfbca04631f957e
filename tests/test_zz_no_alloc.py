"""pg_sweep_noalloc, the entry point that allocates nothing, against the
solver's own sweep (pg_sweep) through the C ABI: the same V bit for bit, the
same sup-norm, the same summed decrease (bit for bit against the
global-gather sweep it shares its kernel with). Two problems in a row, the
second smaller, so the shared scratch is reused with a smaller grid; then a
grid above the scratch's limit is refused before anything runs. Device
buffers come from torch (plumbing only); the library's stream is
synchronised by the entry point before it returns.
"""
import ctypes
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def solver_sweeps(pg, prob, n, kernel):
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        if kernel is not None:
            s.set_kernel(kernel)
        s.fill_seed(prob.seeds)
        V0 = s.download()
        out = []
        for it in range(n):
            r = s.sweep(it, it % 2 == 1)
            out.append((s.download(), r.sum_decrease, r.linf))
        return V0, out


def noalloc(pg, torch, prob, vin, vout, occ, phase):
    ctx = pg.default_context()
    lib = ctx.lib
    gc, mc, pc = prob.grid.c(), prob.model.c_struct(), prob.params.c()
    soa = pg.ctl_soa(prob.ctl)
    d = ctypes.c_double()
    l = ctypes.c_float()
    torch.cuda.synchronize()
    rc = lib.pg_sweep_noalloc(ctx._h, ctypes.byref(gc), ctypes.byref(mc), ctypes.byref(pc), soa.ctypes.data,
                              prob.nctl, occ.data_ptr(), vin.data_ptr(), vout.data_ptr(), int(phase),
                              ctypes.byref(d), ctypes.byref(l))
    return rc, d.value, l.value


def cases(pg):
    big = pg.configs.problem(2, shape=(16, 16, 12, 7, 7, 9))
    small = pg.configs.problem(2, shape=(12, 12, 10, 5, 5, 7))
    # the reference's production candidates (no warm start): global kernel only
    prod = dataclasses.replace(small, params=dataclasses.replace(
        small.params, ntau=5, tau_ratio=2.0, cfl=2.0, tau_min=0.004, tau_max=0.5, rounds=2, delta0=0.35,
        ncoarse=8, simplex=True, rk2=True, adaptive_checks=True),
        model=dataclasses.replace(small.model, knee=0.6))
    return [("problem 2 at 16x16x12x7x7x9", big), ("problem 2 at 12x12x10x5x5x7", small),
            ("production scheme at 12x12x10x5x5x7", prod)]


def test_noalloc_sweeps_match_pg_sweep(pg):
    import torch
    for name, prob in cases(pg):
        print(f"no-alloc sweep: {name}, {prob.grid.ncells} cells", flush=True)
        V0, auto = solver_sweeps(pg, prob, 3, None)
        _, glob = solver_sweeps(pg, prob, 3, pg.abi.PG_KERNEL_GLOBAL)
        occ = torch.from_numpy(np.ascontiguousarray(prob.occ, np.uint8)).cuda()
        a = torch.from_numpy(V0.copy()).cuda()
        b = torch.full_like(a, float("nan"))
        for it in range(3):
            rc, d, l = noalloc(pg, torch, prob, a, b, occ, it)
            assert rc == 0, pg.default_context().lib.pg_last_error()
            got = b.cpu().numpy()
            Vw, dw, lw = auto[it]
            assert bits_equal(got, Vw), f"{name} sweep {it}: {np.count_nonzero(got != Vw)} cells differ"
            assert l == lw
            assert d == pytest.approx(dw, rel=1e-9, abs=1e-6)
            Vg, dg, lg = glob[it]
            assert bits_equal(got, Vg) and l == lg and d == dg
            a, b = b, a
        print(f"  3 sweeps bit-identical, last linf {l:.4f}", flush=True)


def test_noalloc_refuses_grids_above_its_scratch(pg):
    import torch
    prob = pg.configs.problem(2, shape=(12, 12, 10, 5, 5, 7))
    big = pg.configs.make_grid((96, 96, 64, 17, 17, 21))
    assert big.ncells > pg.abi.PG_NOALLOC_MAX_CELLS
    over = dataclasses.replace(prob, grid=big)
    t = torch.zeros(16, device="cuda")
    occ = torch.zeros(16, dtype=torch.uint8, device="cuda")
    rc, _, _ = noalloc(pg, torch, over, t, t[8:], occ, 0)
    assert rc == pg.abi.PG_ERR_ARG
    assert b"at most" in pg.default_context().lib.pg_last_error()
