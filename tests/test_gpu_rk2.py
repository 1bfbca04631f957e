"""Explicit midpoint (RK2) substeps on the device against the CPU oracle.

step_state_rk2 (PeregrineSolver.jl:259-279) is the reference's default
integrator. The exact kernels integrate with it bit for bit as the oracle
restates it (whose order is pinned in test_oracle_pins.py), so the bar is
the plain parity bar: every V value and argmin index equal, sweep by sweep.
"""
import dataclasses

import numpy as np
import pytest

from conftest import oracle_problem
from test_gpu_parity import step_parity

pytestmark = pytest.mark.gpu


def with_rk2(prob):
    return dataclasses.replace(prob, params=dataclasses.replace(prob.params, rk2=True))


@pytest.mark.parametrize("kernel", ["global", "lds"])
def test_config1_rk2_sweeps_bit_exact(pg, orc, kernel):
    """537,600 states, 27 controls, RK2: three sweeps from the seeded start."""
    step_parity(pg, orc, with_rk2(pg.configs.problem(1)), 3, kernel=kernel)


def test_obstacles_125_controls_rk2_bit_exact(pg, orc):
    """A reduced problem-2 grid with obstacles and 125 controls, RK2."""
    prob = with_rk2(pg.configs.problem(2, shape=(16, 16, 12, 7, 7, 9)))
    step_parity(pg, orc, prob, 4, kernel="lds")


def test_rk2_converged_table_is_oracle_fixed_point(pg, orc):
    """A device RK2 solve (active boxes on) stops at a table that two oracle
    sweeps reproduce bit for bit, argmin included; RK2 and Euler tables
    differ (the integrator is really switched)."""
    prob = with_rk2(pg.configs.problem(1))
    tabs = {}
    for p in (prob, pg.configs.problem(1)):
        with pg.HipSolver(p.grid, p.model, p.params, p.ctl, p.occ) as s:
            s.set_kernel(pg.abi.PG_KERNEL_LDS)
            s.fill_seed(p.seeds)
            its, _, linf = s.solve(400, -1.0, 1e-3)
            assert linf < 1e-3 and its < 400
            tabs[p.params.rk2] = s.download()
    step_parity(pg, orc, prob, 2, start=tabs[True], kernel="lds")
    assert not np.array_equal(tabs[True], tabs[False])


def test_rk2_is_refused_where_not_built(pg):
    """The precision form and the FP64 tables integrate with Euler only."""
    prob = with_rk2(pg.configs.problem(1))
    with pytest.raises(pg.abi.PeregrineError):
        pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ, vstore=pg.abi.PG_VSTORE_U16_TILE)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        with pytest.raises(pg.abi.PeregrineError):
            s.set_kernel(pg.abi.PG_KERNEL_FAST)
        with pytest.raises(pg.abi.PeregrineError):
            s.solve_f64(1, 0.0)
