"""Out-of-core sweeps against the in-HBM sweep of the same grid.

The tiled sweep streams each tile's window of the old table from pinned host
memory, sweeps the tile's interior and streams it back. Jacobi sweeps read
only the old table, so the bar is bit equality of every value after every
sweep, and the same ||dV||_inf (the summed decrease is added per tile, so it
agrees to rounding only).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def compare_sweeps(pg, prob, sweeps, budget, tile):
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as ref, \
         pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ,
                      ooc_budget=budget, ooc_tile=tile) as ooc:
        ref.set_kernel(pg.abi.PG_KERNEL_LDS)
        ref.fill_seed(prob.seeds)
        ooc.fill_seed(prob.seeds)
        info = ooc.ooc_info()
        print(f"tiles {info[0]} x {info[1]}, {info[2] / 2**20:.1f} MiB on the device, "
              f"amplification {info[3]:.2f}")
        for it in range(sweeps):
            r = ref.sweep(it, it % 2 == 1)
            o = ooc.sweep(it, it % 2 == 1)
            a, b = ref.download(), ooc.download()
            bad = np.nonzero(bits(a) != bits(b))[0]
            assert bad.size == 0, f"sweep {it}: {bad.size} cells differ, first {bad[:5].tolist()}"
            assert o.linf == r.linf
            assert o.sum_decrease == pytest.approx(r.sum_decrease, rel=1e-9, abs=1e-6)
        return info


def test_tiles_cut_x_y_heading(pg):
    """36 tiles of 8 x 8 x 4 interior cells (heading windows wrap) on a
    reduced problem-2 grid with obstacles and 125 controls, 6 sweeps."""
    prob = pg.configs.problem(2, shape=(24, 24, 16, 9, 9, 11))
    info = compare_sweeps(pg, prob, 6, 8 << 30, (8, 8, 4))
    assert info[0] == (8, 8, 4) and info[1] == 36


def test_tiles_with_partial_edges(pg):
    """A tile request off the box edge (5, 7, 3) rounds up to (6, 8, 4);
    the last tile on each axis is partial."""
    prob = pg.configs.problem(1)
    info = compare_sweeps(pg, prob, 5, 8 << 30, (5, 7, 3))
    assert info[0] == (6, 8, 4)


def test_tiles_from_budget(pg):
    """With no tile given, the largest x-y tile (whole heading axis) whose
    buffers fit the budget; a budget far below one table forces several."""
    prob = pg.configs.problem(2, shape=(32, 32, 16, 9, 9, 11))
    table_bytes = prob.grid.ncells * 4
    info = compare_sweeps(pg, prob, 4, table_bytes // 2, (0, 0, 0))
    assert info[1] > 1 and info[2] <= table_bytes // 2
    assert info[0][2] == 16


def test_solve_matches_in_hbm_solve(pg):
    """A whole solve to ||dV||_inf < 1e-3: same sweep count, same table."""
    prob = pg.configs.problem(1)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as ref, \
         pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ,
                      ooc_budget=1 << 30, ooc_tile=(4, 4, 6)) as ooc:
        ref.set_kernel(pg.abi.PG_KERNEL_LDS)
        ref.fill_seed(prob.seeds)
        ooc.fill_seed(prob.seeds)
        its_r, _, l_r = ref.solve(2000, -1.0, 1e-3)
        its_o, _, l_o = ooc.solve(2000, -1.0, 1e-3)
        assert its_r == its_o and l_r == l_o
        assert np.array_equal(bits(ref.download()), bits(ooc.download()))


def test_budget_too_small_is_an_error(pg):
    prob = pg.configs.problem(1)
    with pytest.raises(pg.abi.PeregrineError):
        pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ, ooc_budget=1 << 16,
                     ooc_tile=(0, 0, 0))


@pytest.mark.parametrize("store", ["u16", "f16"])
def test_precision_form_out_of_core(pg, store):
    """The precision form out of core (codes streamed, offsets/scales refit
    tile by tile into the next sweep's array) against the in-HBM precision
    form: 6 sweeps with x, y and wrapping heading cuts, each bit for bit,
    then 2 FP32 polish sweeps."""
    vs = pg.abi.PG_VSTORE_U16_TILE if store == "u16" else pg.abi.PG_VSTORE_F16_TILE
    prob = pg.configs.problem(2, shape=(24, 24, 16, 9, 9, 11))
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ, vstore=vs) as ref, \
         pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ, vstore=vs,
                      ooc_budget=8 << 30, ooc_tile=(8, 8, 8)) as ooc:
        ref.fill_seed(prob.seeds)
        ooc.fill_seed(prob.seeds)
        info = ooc.ooc_info()
        print(f"{store}: tiles {info[0]} x {info[1]}, amplification {info[3]:.2f}", flush=True)
        assert info[0] == (8, 8, 8) and info[1] == 18
        for it in range(6):
            r = ref.sweep(it, it % 2 == 1)
            o = ooc.sweep(it, it % 2 == 1)
            a, b = ref.download(), ooc.download()
            bad = np.nonzero(bits(a) != bits(b))[0]
            assert bad.size == 0, f"sweep {it}: {bad.size} cells differ, first {bad[:5].tolist()}"
            assert o.linf == r.linf
            assert o.sum_decrease == pytest.approx(r.sum_decrease, rel=1e-9, abs=1e-6)
        ra = ref.polish(2)
        oa = ooc.polish(2)
        assert np.array_equal(bits(ref.download()), bits(ooc.download()))
        assert ra[1] == oa[1]


@pytest.mark.parametrize("shape,tile", [((24, 24, 16, 9, 9, 11), (8, 8, 4)), (None, (5, 7, 3))])
def test_escape_pass_out_of_core(pg, shape, tile):
    """The escape pass out of core (windows planned for the escape reads:
    blocked cells and probe-rejected successors included) against the
    in-HBM escape pass: 4 value sweeps, then 4 escape sweeps, each bit for
    bit, on a reduced problem-2 grid with wrapping heading windows and on
    problem 1 with partial edge tiles."""
    prob = pg.configs.problem(2, shape=shape) if shape else pg.configs.problem(1)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as ref, \
         pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ,
                      ooc_budget=8 << 30, ooc_tile=tile) as ooc:
        ref.fill_seed(prob.seeds)
        ooc.fill_seed(prob.seeds)
        print(f"escape: grid {prob.grid.n}, tiles {ooc.ooc_info()[:2]}", flush=True)
        for it in range(4):
            ref.sweep(it, it % 2 == 1)
            ooc.sweep(it, it % 2 == 1)
        assert np.array_equal(bits(ref.download()), bits(ooc.download()))
        for it in range(4):
            r = ref.escape_sweep(it, it % 2 == 1)
            o = ooc.escape_sweep(it, it % 2 == 1)
            a, b = ref.download(), ooc.download()
            bad = np.nonzero(bits(a) != bits(b))[0]
            assert bad.size == 0, f"escape sweep {it}: {bad.size} cells differ, first {bad[:5].tolist()}"
            assert o.linf == r.linf
            assert o.sum_decrease == pytest.approx(r.sum_decrease, rel=1e-9, abs=1e-6)
        V = ooc.download()
        assert (V < 0).sum() > 0


def test_production_scheme_out_of_core(pg):
    """The reference's default scheme out of core (pattern search, horizon
    ladder, Int8 warm start, Kuhn lookups, RK2, CFL horizon, adaptive probes,
    rotating slice; windows from the scheme's reach, warm-start entries
    streamed beside each interior) against the in-HBM global-gather kernel:
    5 value sweeps, V and the warm-start store bit for bit, then 2 escape
    sweeps with the production candidates."""
    import dataclasses
    prob = pg.configs.problem(2, shape=(24, 24, 16, 5, 5, 7))
    prob = dataclasses.replace(prob, model=dataclasses.replace(prob.model, knee=0.6))
    prob = dataclasses.replace(prob, params=dataclasses.replace(
        prob.params, ntau=5, tau_ratio=2.0, cfl=2.0, tau_min=0.004, tau_max=0.5, rounds=2, delta0=0.35, ncoarse=8,
        simplex=True, rk2=True, adaptive_checks=True))
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as ref, \
         pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ,
                      ooc_budget=8 << 30, ooc_tile=(8, 8, 4)) as ooc:
        ref.set_warm_start(True)
        ooc.set_warm_start(True)
        ref.fill_seed(prob.seeds)
        ooc.fill_seed(prob.seeds)
        print(f"production out of core: grid {prob.grid.n}, tiles {ooc.ooc_info()[:2]}", flush=True)
        for it in range(5):
            r = ref.sweep(it, it % 2 == 1)
            o = ooc.sweep(it, it % 2 == 1)
            a, b = ref.download(), ooc.download()
            bad = np.nonzero(bits(a) != bits(b))[0]
            assert bad.size == 0, f"sweep {it}: {bad.size} cells differ, first {bad[:5].tolist()}"
            assert o.linf == r.linf
            assert o.sum_decrease == pytest.approx(r.sum_decrease, rel=1e-9, abs=1e-6)
            assert np.array_equal(ref.download_warm_start(), ooc.download_warm_start()), f"sweep {it}: store differs"
        assert np.count_nonzero(ooc.download_warm_start()) > 0
        for it in range(2):
            r = ref.escape_sweep(it, it % 2 == 1)
            o = ooc.escape_sweep(it, it % 2 == 1)
            assert np.array_equal(bits(ref.download()), bits(ooc.download())), f"escape sweep {it} differs"
            assert o.linf == r.linf
