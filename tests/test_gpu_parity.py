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


KERNELS = ["global", "lds"]


def kernel_id(pg, name):
    return {"global": pg.abi.PG_KERNEL_GLOBAL, "lds": pg.abi.PG_KERNEL_LDS}[name]


def step_parity(pg, orc, prob, sweeps, start=None, kernel="lds"):
    op = oracle_problem(orc, prob)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.set_kernel(kernel_id(pg, kernel))
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


@pytest.mark.parametrize("kernel", KERNELS)
def test_config1_sweeps_bit_exact(pg, orc, kernel):
    """537,600 states, 27 controls: three sweeps from the seeded start."""
    prob = pg.configs.problem(1)
    step_parity(pg, orc, prob, 3, kernel=kernel)


@pytest.mark.parametrize("kernel", KERNELS)
def test_config1_converged_table_bit_exact(pg, orc, kernel):
    """A converged table (GPU solve) is a fixed point of the oracle's sweep
    too, and its argmin policy matches."""
    prob = pg.configs.problem(1)
    op = oracle_problem(orc, prob)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.set_kernel(kernel_id(pg, kernel))
        s.fill_seed(prob.seeds)
        its, _, linf = s.solve(400, -1.0, 1e-3)
        V = s.download()
    assert linf < 1e-3 and its < 400
    step_parity(pg, orc, prob, 2, start=V, kernel=kernel)
    reached = np.count_nonzero(V < prob.params.cap) / V.size
    assert reached > 0.5


@pytest.mark.parametrize("kernel", KERNELS)
def test_obstacles_125_controls_bit_exact(pg, orc, kernel):
    """Problem 2's model, obstacles and 125 controls on a reduced grid, from
    a partly propagated table (so obstacles and cap boundaries are active)."""
    prob = pg.configs.problem(2, shape=(16, 16, 12, 7, 7, 9))
    assert prob.occ.any()
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        s.solve(12, -1.0, 0.0)
        V = s.download()
    step_parity(pg, orc, prob, 2, start=V, kernel=kernel)


def test_plan_probe_flags(pg):
    """The exact plan's per-box probe flags (pg_solver_probe_info), which let
    the LDS-staged sweeps skip occupancy lookups or whole probe loops: with no
    obstacle every box's chords are free and the interior boxes skip their
    probes; with problem 2's obstacles some boxes keep their lookups. The
    sweeps these flags shape are the bit-exact ones of the tests above."""
    def flags(prob):
        with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
            s.set_kernel(pg.abi.PG_KERNEL_LDS)
            assert s.probe_info() == (0.0, 0.0)      # before the plan
            s.fill_seed(prob.seeds)
            s.sweep(0, False)
            return s.probe_info()

    free1, pass1 = flags(pg.configs.problem(1))
    print(f"problem 1: chords free {free1:.3f}, probes skipped {pass1:.3f}", flush=True)
    assert free1 == 1.0 and 0.0 < pass1 < 1.0
    free2, pass2 = flags(pg.configs.problem(2, shape=(16, 16, 12, 7, 7, 9)))
    print(f"problem 2 (reduced): chords free {free2:.3f}, probes skipped {pass2:.3f}", flush=True)
    assert 0.0 < free2 < 1.0 and 0.0 <= pass2 < 1.0


@pytest.mark.parametrize("kernel", KERNELS)
def test_full_solve_matches_oracle_solve(pg, orc, kernel):
    """Whole solve on a small grid: same iteration count, same table."""
    prob = pg.configs.problem(1, shape=(10, 10, 8, 5, 5, 5))
    op = oracle_problem(orc, prob)
    Vo, ito, lo = op.solve(prob.seeds, 300, 1e-3, nthreads=nthreads())
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.set_kernel(kernel_id(pg, kernel))
        s.fill_seed(prob.seeds)
        its, _, linf = s.solve(300, -1.0, 1e-3)
        Vg = s.download()
    assert its == ito
    assert bits_equal(Vg, Vo), diff_report(Vg, Vo)


def realistic_model(pg):
    """The reference's `_realistic()` fit (SelfTest.jl:51-61): dense blocks,
    traction knee 0.45 -- the kernels' general (non-diagonal) path."""
    B = (-849.0, -25.8, -69.5, 5.0, -846.8, 0.6, -0.2, -9.6, -64.9)
    A = (-2.13, -0.007, -8.72, -0.064, -4.96, -0.013, 0.039, 0.078, -4.67)
    S = (-11.6, -3.37, 3.29, 2.07, 6.33, 1.00, -0.56, -1.17, -0.28)
    D = (0.021, 0.003, 0.729, 0.001, 0.049, 0.871, -0.0004, -0.0006, 0.312)
    return pg.Model(B, A, (0.0, 0.0, 0.0), S, D, (0.0, 0.0, 0.0), (5.0, 5.0, 0.15), 0.45)


@pytest.mark.parametrize("kernel", KERNELS)
def test_dense_model_with_knee_bit_exact(pg, orc, kernel):
    import dataclasses
    prob = pg.configs.problem(2, shape=(12, 12, 8, 7, 7, 9))
    prob = dataclasses.replace(prob, model=realistic_model(pg), ctl=pg.control_set(4))
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        s.solve(10, -1.0, 0.0)
        V = s.download()
    step_parity(pg, orc, prob, 2, start=V, kernel=kernel)


def test_full_size_problem2_lds_equals_global_and_oracle(pg, orc):
    """564,854,784 states x 125 controls: the LDS-staged sweep writes the
    same table as the global-gather sweep, and random cells match the
    oracle's backup of the same V."""
    prob = pg.configs.problem(2)
    op = oracle_problem(orc, prob)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.set_kernel(pg.abi.PG_KERNEL_LDS)
        s.fill_seed(prob.seeds)
        for _ in range(3):
            s.sweep()
        V3 = s.download()
        r_lds = s.sweep()
        V4_lds = s.download()
        kern, cov = s.kernel_info()
        assert kern == pg.abi.PG_KERNEL_LDS and cov > 0.9
        s.upload(V3)
        s.set_kernel(pg.abi.PG_KERNEL_GLOBAL)
        r_glb = s.sweep()
        V4_glb = s.download()
    assert bits_equal(V4_lds, V4_glb), diff_report(V4_lds, V4_glb)
    assert r_lds.linf == r_glb.linf
    rng = np.random.default_rng(17)
    cells = rng.integers(0, prob.grid.ncells, 20000)
    c = op.backup_cells_f32(V3, cells, nthreads())
    want = np.minimum(c, V3[cells])
    want[np.isin(cells, prob.seeds)] = 0.0
    assert bits_equal(V4_lds[cells], want), diff_report(V4_lds[cells], want)


def test_box_grid_in_two_launch_chunks(pg):
    """A grid of 11.9M boxes is swept in two launches of at most 2^23
    workgroups (a dispatch counts work-items in 32 bits): one LDS sweep of a
    random table equals the global-gather sweep bit for bit, so every box of
    both chunks was swept once."""
    prob = pg.configs.problem(2, shape=(64, 64, 48, 17, 17, 21))
    rng = np.random.default_rng(7)
    V0 = rng.uniform(0.0, 61.0, prob.grid.ncells).astype(np.float32)
    V0[V0 > 60.0] = 60.0
    V0[prob.seeds] = 0.0
    out = {}
    for kern in ("global", "lds"):
        with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
            s.set_kernel(kernel_id(pg, kern))
            s.upload(V0)
            s.sweep(0, False)
            out[kern] = s.download()
    assert bits_equal(out["global"], out["lds"]), diff_report(out["lds"], out["global"])
    assert np.count_nonzero(out["lds"] != V0) > 0.5 * V0.size
