"""The rest of the reference's production scheme on the device against the
CPU oracle: pattern search (refine_rounds), the horizon ladder
(tau_levels), the Int8 warm-start store, and all of them together with the
options already built (the reference's default solve.jl scheme). The bar is
the parity bar: V bit-identical sweep by sweep, and the warm-start store
byte-identical. A traction knee is set so that commands the device builds
itself (search moves, stored commands) go through its own traction gain.
"""
import dataclasses

import numpy as np
import pytest

from conftest import nthreads, oracle_problem
from test_gpu_parity import bits_equal, diff_report

pytestmark = pytest.mark.gpu

SHAPE = (12, 12, 10, 5, 5, 7)
CASES = {
    "search": (dict(rounds=2, delta0=0.35), False),
    "ladder": (dict(ntau=5, tau_ratio=2.0, cfl=2.0, tau_min=0.004, tau_max=0.5), False),
    "warm": (dict(), True),
    "reference_default": (dict(ntau=5, tau_ratio=2.0, cfl=2.0, tau_min=0.004, tau_max=0.5, rounds=2, delta0=0.35,
                               ncoarse=8, simplex=True, rk2=True, adaptive_checks=True), True),
}


def problem(pg, **kw):
    prob = pg.configs.problem(2, shape=SHAPE)
    prob = dataclasses.replace(prob, model=dataclasses.replace(prob.model, knee=0.6))
    return dataclasses.replace(prob, params=dataclasses.replace(prob.params, **kw))


@pytest.mark.parametrize("case", list(CASES))
def test_production_sweeps_bit_exact(pg, orc, case):
    kw, warm = CASES[case]
    prob = problem(pg, **kw)
    op = oracle_problem(orc, prob)
    n = prob.grid.ncells
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        if warm:
            s.set_warm_start(True)
        kern = pg.abi.PG_KERNEL_GLOBAL
        assert s.kernel_info()[0] == kern
        s.fill_seed(prob.seeds)
        V = s.download()
        wp = np.zeros(3 * n, np.int8)
        for it in range(5):
            r = s.sweep(it, it % 2 == 1)
            got = s.download()
            want = np.empty_like(V)
            if warm:
                d, l = op.sweep_warm(V, want, wp, phase=it, nthreads=nthreads())
            else:
                d, l = op.sweep(V, want, phase=it, nthreads=nthreads())
            want[prob.seeds] = 0.0
            assert bits_equal(got, want), f"{case} sweep {it}: " + diff_report(got, want)
            assert r.linf == l
            assert r.sum_decrease == pytest.approx(d, rel=1e-9, abs=1e-6)
            if warm:
                gw = s.download_warm_start()
                assert np.array_equal(gw, wp), f"{case} sweep {it}: {np.count_nonzero(gw != wp)} store bytes differ"
            V = got
        if warm:
            assert np.count_nonzero(wp) > 0


def test_production_escape_bit_exact(pg, orc):
    """The escape pass runs the same search and ladder (no warm start)."""
    prob = problem(pg, rounds=2, delta0=0.35, ntau=3, cfl=2.0, tau_min=0.004, tau_max=0.5, ncoarse=8)
    op = oracle_problem(orc, prob)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        for it in range(20):
            s.sweep(it)
        V = s.download()
        for it in range(4):
            r = s.escape_sweep(it, it % 2 == 1)
            got = s.download()
            want = np.empty_like(V)
            d, l = op.sweep(V, want, phase=it, nthreads=nthreads(), esc=1)
            assert bits_equal(got, want), f"escape sweep {it}: " + diff_report(got, want)
            assert r.linf == l
            V = got
        assert np.count_nonzero(V < 0) > 0


def test_production_options_refused_elsewhere(pg):
    """Kernels that cannot evaluate commands or horizons unknown before the
    sweep refuse them by name."""
    prob = problem(pg, rounds=2)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        with pytest.raises(pg.abi.PeregrineError, match="refine_rounds"):
            s.set_kernel(pg.abi.PG_KERNEL_LDS)
    with pytest.raises(pg.abi.PeregrineError):
        pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ, vstore=pg.abi.PG_VSTORE_U16_TILE)
    # out of core the FP32 store runs them (tests/test_gpu_ooc.py), the precision form does not
    with pytest.raises(pg.abi.PeregrineError, match="out-of-core precision form"):
        pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ, ooc_budget=8 << 30,
                     vstore=pg.abi.PG_VSTORE_U16_TILE)
    plain = problem(pg)
    with pg.HipSolver(plain.grid, plain.model, plain.params, plain.ctl, plain.occ) as s:
        s.set_warm_start(True)
        with pytest.raises(pg.abi.PeregrineError, match="warm_start"):
            s.set_kernel(pg.abi.PG_KERNEL_LDS)
        s.set_warm_start(False)
        s.set_kernel(pg.abi.PG_KERNEL_LDS)
