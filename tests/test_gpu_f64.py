"""GPU FP64 tables against the CPU FP64 oracle.

The FP64 sweeps carry the plain scheme in double on the device; the oracle
carries the same algorithm in double on the CPU (oracle/pg_oracle.hpp with
T = double). The two differ only through the sin/cos/tanh of the device and
host libraries (a few ulp of double), so the bar is a small absolute
tolerance, written in each test, not bit equality.
"""
import numpy as np
import pytest

from conftest import nthreads, oracle_problem

pytestmark = pytest.mark.gpu

SMALL = (8, 8, 6, 5, 5, 5)     # 9,600 states, 27 controls


def test_f64_plain_sweeps_match_oracle(pg, orc):
    """Four FP64 Jacobi sweeps from the seeded start, device vs oracle."""
    prob = pg.configs.problem(1, shape=SMALL)
    op = oracle_problem(orc, prob)
    cap = prob.params.cap
    V = np.full(prob.grid.ncells, cap, np.float64)
    V[prob.seeds] = 0.0
    W = np.empty_like(V)
    for _ in range(4):
        op.sweep(V, W, nthreads=nthreads())
        W[prob.seeds] = 0.0
        V, W = W, V
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        its, _ = s.solve_f64(4, 0.0, diagonal=False)
        got = s.download_f64()
    assert its == 4
    assert np.array_equal(got < cap, V < cap)
    assert np.max(np.abs(got - V)) < 1e-9


def test_f64_fixed_point_certified_by_oracle(pg, orc):
    """The diagonal-solved FP64 table is a fixed point of the oracle's own
    FP64 backup: one CPU FP64 backup of every state but the targets differs
    from the table by at most 1e-9 s, and the device's residual pass agrees."""
    prob = pg.configs.problem(1, shape=SMALL)
    op = oracle_problem(orc, prob)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        its, l = s.solve_f64(3000, 1e-12)
        V = s.download_f64()
        r_max, _ = s.residual_f64(0)
        pol = s.policy_f64()
    assert l < 1e-12, (its, l)
    assert r_max < 1e-9
    TV, kk = op.backup_cells_f64(V, np.arange(prob.grid.ncells), nthreads=nthreads())
    cap = prob.params.cap
    live = (V < cap) | (TV < cap)
    live[prob.seeds] = False          # targets are held at 0, not backed up
    assert np.max(np.abs(TV - V)[live]) < 1e-9
    # argmin agreement away from near-ties
    assert np.mean(kk.astype(np.uint8) == pol) > 0.999


def test_f64_diagonal_and_plain_reach_the_same_table(pg):
    """Both FP64 iterations converge to one fixed point; the diagonal solve
    gets there in fewer sweeps."""
    prob = pg.configs.problem(1, shape=(12, 12, 8, 5, 5, 7))
    tabs = {}
    for diagonal in (False, True):
        with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
            s.fill_seed(prob.seeds)
            its, l = s.solve_f64(5000, 1e-12, diagonal=diagonal)
            assert l < 1e-12
            tabs[diagonal] = (s.download_f64(), its)
    print(f"FP64 sweeps to 1e-12: plain {tabs[False][1]}, diagonal {tabs[True][1]}")
    assert tabs[True][1] < tabs[False][1]
    assert np.max(np.abs(tabs[True][0] - tabs[False][0])) < 1e-9


def test_fp32_table_against_fp64(pg):
    """The exact FP32 table (the reference's arithmetic), solved until it
    stops changing, against the FP64 table on a reduced problem-2 grid
    (obstacles, 125 controls); the device FP64 residual of the FP32 table.
    Recorded: the FP32 table's own distance from FP64 is the floor for any
    faster form. Guards: p99 below 1e-3 s, reachability equal."""
    prob = pg.configs.problem(2, shape=(24, 24, 16, 9, 9, 11))
    cap = prob.params.cap
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        its32, d, _ = s.solve(4000, 0.0, 0.0)
        V32 = s.download().astype(np.float64)
        its64, l64 = s.solve_f64(3000, 1e-10)
        V64 = s.download_f64()
        r64, _ = s.residual_f64(0)
        r32, _ = s.residual_f64(1)
    both = (V32 < cap) & (V64 < cap)
    a = np.abs(V32 - V64)[both]
    q = np.quantile(a, [0.5, 0.99, 0.999])
    print(f"FP32 ({its32} sweeps) vs FP64 ({its64} diagonal sweeps, linf {l64:.1e}): max {a.max():.2e} "
          f"p50 {q[0]:.1e} p99 {q[1]:.1e} p99.9 {q[2]:.1e} mean signed {np.mean((V32 - V64)[both]):+.1e}; "
          f"FP64 residual of the FP64 table {r64:.1e}, of the FP32 table {r32:.1e}; "
          f"reach mismatch {np.mean((V32 < cap) != (V64 < cap)):.1e}")
    assert l64 < 1e-10
    assert np.mean((V32 < cap) != (V64 < cap)) < 1e-4
    assert q[1] < 1e-3
