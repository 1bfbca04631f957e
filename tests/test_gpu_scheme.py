"""The scheme's options on the device against the CPU oracle: nearest-node
lookups, velocity_clamp, adaptive probe counts, the rotating control slice
(control_scan), Kuhn simplex lookups and the per-cell CFL horizon, alone and
together, in both exact kernels. The bar is the parity bar: every V value equal, sweep by sweep,
each sweep at its own phase (the slice rotates with it).
"""
import dataclasses

import numpy as np
import pytest

from conftest import nthreads, oracle_problem
from test_gpu_parity import bits_equal, diff_report

pytestmark = pytest.mark.gpu

OPTIONS = {
    "nearest": dict(nearest=True),
    "vclamp": dict(vclamp=True),
    "adaptive": dict(adaptive_checks=True),
    "slice": dict(ncoarse=24),
    "all": dict(nearest=True, vclamp=True, adaptive_checks=True, ncoarse=24, rk2=True),
    "simplex": dict(simplex=True),
    "cfl": dict(cfl=0.5),
    "production": dict(simplex=True, cfl=0.5, rk2=True, adaptive_checks=True, ncoarse=24, vclamp=True),
}


def with_params(prob, **kw):
    return dataclasses.replace(prob, params=dataclasses.replace(prob.params, **kw))


def phased_parity(pg, orc, prob, sweeps, kernel):
    op = oracle_problem(orc, prob)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.set_kernel(kernel)
        s.fill_seed(prob.seeds)
        V = s.download()
        for it in range(sweeps):
            r = s.sweep(it, it % 2 == 1)
            got = s.download()
            want = np.empty_like(V)
            d, l = op.sweep(V, want, phase=it, nthreads=nthreads())
            want[prob.seeds] = 0.0
            assert bits_equal(got, want), f"sweep {it}: " + diff_report(got, want)
            assert r.linf == l
            assert r.sum_decrease == pytest.approx(d, rel=1e-9, abs=1e-6)
            V = got
        return s.policy(), V


@pytest.mark.parametrize("kernel", ["global", "lds"])
@pytest.mark.parametrize("option", list(OPTIONS))
def test_option_sweeps_bit_exact(pg, orc, option, kernel):
    """A reduced problem-2 grid (obstacles, 125 controls), 4 sweeps."""
    prob = with_params(pg.configs.problem(2, shape=(16, 16, 12, 7, 7, 9)), **OPTIONS[option])
    kid = pg.abi.PG_KERNEL_GLOBAL if kernel == "global" else pg.abi.PG_KERNEL_LDS
    pol, V = phased_parity(pg, orc, prob, 4, kid)
    # the policy is the argmin over the whole lattice
    full = with_params(prob, ncoarse=0)
    want_pol = np.empty(prob.grid.ncells, np.uint8)
    oracle_problem(orc, full).sweep(V, np.empty_like(V), pol=want_pol, nthreads=nthreads())
    assert np.array_equal(pol, want_pol), f"{np.count_nonzero(pol != want_pol)} policy cells differ"


def test_slice_solve_needs_a_quiet_turn(pg):
    """With a slice of 24 of 124 moving controls, pg_solve stops only after
    ceil(124/24) = 6 quiet sweeps in a row; the stopped table is a fixed
    point of the full-lattice sweep to the tolerance."""
    prob = with_params(pg.configs.problem(2, shape=(16, 16, 12, 7, 7, 9)), ncoarse=24)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        its, _, l = s.solve(3000, -1.0, 1e-3)
        assert l < 1e-3 and its < 3000
        V = s.download()
    full = with_params(prob, ncoarse=0)
    with pg.HipSolver(full.grid, full.model, full.params, full.ctl, full.occ) as s:
        s.fill_seed(full.seeds)
        s.upload(V)
        r = s.sweep(0)
    assert r.linf < 1e-3


def test_options_refused_by_the_precision_form(pg):
    prob = with_params(pg.configs.problem(1), nearest=True)
    with pytest.raises(pg.abi.PeregrineError):
        pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ, vstore=pg.abi.PG_VSTORE_U16_TILE)


def test_slice_out_of_core_equals_in_hbm(pg):
    """The tiled sweep carries the phase's slice too: bit-identical to in-HBM."""
    prob = with_params(pg.configs.problem(2, shape=(24, 24, 16, 9, 9, 11)), ncoarse=24, adaptive_checks=True)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as ref, \
         pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ, ooc_budget=8 << 30,
                      ooc_tile=(8, 8, 4)) as ooc:
        ref.set_kernel(pg.abi.PG_KERNEL_LDS)
        ref.fill_seed(prob.seeds)
        ooc.fill_seed(prob.seeds)
        for it in range(5):
            a = ref.sweep(it)
            b = ooc.sweep(it)
            assert bits_equal(ref.download(), ooc.download()), f"sweep {it}"
            assert a.linf == b.linf
