"""GPU parity: the HIP sweep against the CPU oracle, through the C ABI.

The bar is bit equality of every V value and every argmin index: the plain
FP32 kernel reproduces the reference's unfused float arithmetic exactly, and
a Jacobi sweep has no scheduling freedom. The oracle steps from the GPU's
own V each time, so one mismatch pins the sweep and cell where it appeared.
"""
import numpy as np
import pytest

from conftest import nthreads, oracle_problem

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def diff_report(got, want):
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    return f"{bad.size} cells differ; first {bad[:6].tolist()} got {got[bad[:6]].tolist()} want {want[bad[:6]].tolist()}"


def step_parity(pg, orc, prob, sweeps, start=None):
    op = oracle_problem(orc, prob)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        if start is not None:
            s.upload(start)
        V = s.download()
        for it in range(sweeps):
            r = s.sweep(it, it % 2 == 1)
            got = s.download()
            want = np.empty_like(V)
            d, l = op.sweep(V, want, nthreads=nthreads())
            want[prob.seeds] = 0.0
            assert bits_equal(got, want), f"sweep {it}: " + diff_report(got, want)
            assert r.linf == pytest.approx(l, rel=0, abs=0)
            assert r.sum_decrease == pytest.approx(d, rel=1e-9, abs=1e-6)
            V = got
        pol = s.policy()
        want_pol = np.empty(prob.grid.ncells, np.uint8)
        op.sweep(V, np.empty_like(V), pol=want_pol, nthreads=nthreads())
        assert np.array_equal(pol, want_pol), f"{np.count_nonzero(pol != want_pol)} policy cells differ"
        return V, pol


def test_config1_sweeps_bit_exact(pg, orc):
    """537,600 states, 27 controls: three sweeps from the seeded start."""
    prob = pg.configs.problem(1)
    step_parity(pg, orc, prob, 3)


def test_config1_converged_table_bit_exact(pg, orc):
    """A converged table (GPU solve) is a fixed point of the oracle's sweep
    too, and its argmin policy matches."""
    prob = pg.configs.problem(1)
    op = oracle_problem(orc, prob)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        its, _, linf = s.solve(400, -1.0, 1e-3)
        V = s.download()
    assert linf < 1e-3 and its < 400
    step_parity(pg, orc, prob, 2, start=V)
    reached = np.count_nonzero(V < prob.params.cap) / V.size
    assert reached > 0.5


def test_obstacles_125_controls_bit_exact(pg, orc):
    """Problem 2's model, obstacles and 125 controls on a reduced grid, from
    a partly propagated table (so obstacles and cap boundaries are active)."""
    prob = pg.configs.problem(2, shape=(16, 16, 12, 7, 7, 9))
    assert prob.occ.any()
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        s.solve(12, -1.0, 0.0)
        V = s.download()
    step_parity(pg, orc, prob, 2, start=V)


def test_full_solve_matches_oracle_solve(pg, orc):
    """Whole solve on a small grid: same iteration count, same table."""
    prob = pg.configs.problem(1, shape=(10, 10, 8, 5, 5, 5))
    op = oracle_problem(orc, prob)
    Vo, ito, lo = op.solve(prob.seeds, 300, 1e-3, nthreads=nthreads())
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        its, _, linf = s.solve(300, -1.0, 1e-3)
        Vg = s.download()
    assert its == ito
    assert bits_equal(Vg, Vo), diff_report(Vg, Vo)
