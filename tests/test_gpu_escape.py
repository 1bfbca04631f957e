"""The escape pass on the device against the CPU oracle, through the C ABI.

After a value solve, escape sweeps (gpu_escape_kernel!, Gpu.jl:51-67, in
Jacobi form) must equal the oracle's escape sweeps bit for bit, each
stepped from the device's own table; the summed and largest improvements
agree too.
"""
import numpy as np
import pytest

from conftest import nthreads, oracle_problem
from test_gpu_parity import bits_equal, diff_report

pytestmark = pytest.mark.gpu


OPTIONS = dict(rk2=True, simplex=True, adaptive_checks=True, cfl=2.0, tau_min=0.004, tau_max=0.5, ncoarse=16)


@pytest.mark.parametrize("kernel,options", [("lds", False), ("global", False), ("lds", True)])
def test_escape_sweeps_bit_exact(pg, orc, kernel, options):
    """A reduced problem-2 grid (six obstacles, 125 controls): 30 value sweeps,
    then 8 escape sweeps compared one by one; the LDS-staged escape kernel
    and the global-gather one, and the LDS-staged one with the scheme
    options its windows can plan for (RK2, Kuhn lookups, adaptive probes,
    the per-cell horizon, a rotating slice)."""
    import dataclasses
    prob = pg.configs.problem(2, shape=(16, 16, 12, 7, 7, 9))
    if options:
        prob = dataclasses.replace(prob, params=dataclasses.replace(prob.params, **OPTIONS))
    op = oracle_problem(orc, prob)
    cap = np.float32(prob.params.cap)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.set_kernel(pg.abi.PG_KERNEL_LDS if kernel == "lds" else pg.abi.PG_KERNEL_GLOBAL)
        s.fill_seed(prob.seeds)
        s.solve(30, -1.0, 0.0)
        V = s.download()
        routes = (V >= 0) & (V < cap)
        written = 0
        for it in range(8):
            r = s.escape_sweep(it, it % 2 == 1)
            got = s.download()
            want = np.empty_like(V)
            d, l = op.sweep(V, want, phase=it, nthreads=nthreads(), esc=1)
            assert bits_equal(got, want), f"escape sweep {it}: " + diff_report(got, want)
            assert r.linf == l
            assert r.sum_decrease == pytest.approx(d, rel=1e-9, abs=1e-6)
            assert np.array_equal(got[routes], V[routes])
            written = np.count_nonzero(got < 0)
            V = got
        assert written > 0


def test_solve_escape_stops_quiet(pg):
    """pg_solve_escape runs to a quiet sweep; escape times lie in [dt, cap)
    and blocked cells hold some of them."""
    prob = pg.configs.problem(2, shape=(16, 16, 12, 7, 7, 9))
    cap = np.float32(prob.params.cap)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        s.solve(400, -1.0, 1e-3)
        its, d, l = s.solve_escape(400, -1.0, 1e-3)
        assert l < 1e-3 and its < 400
        E = s.download()
    esc = E[E < 0]
    assert esc.size > 0 and np.all(-esc >= prob.params.dt) and np.all(-esc < cap)
    blk = prob.grid.n[3] * prob.grid.n[4] * prob.grid.n[5]
    occ_cells = np.repeat(np.asarray(prob.occ, np.uint8) != 0, blk)
    assert np.count_nonzero(E[occ_cells] < 0) > 0
