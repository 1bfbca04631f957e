"""GPU tests of the precision form: a 16-bit mirror of V (per-4^6-tile FP32
offset/scale) read by the gathers, each state's own value in FP32,
contracted FP32 backups with the diagonal solved, exact FP32 polish.

These are tolerance tests by nature: the fast sweep's arithmetic differs
from the reference's unfused float arithmetic by rounding, and the 16-bit
mirror by its quantization. The tolerances are written in each test.
"""
import numpy as np
import pytest

from conftest import nthreads, oracle_problem

pytestmark = pytest.mark.gpu


TILE = 4   # quantization tile edge of the 16-bit stores


def tile_ranges(g, V, cap):
    """Per-cell (max - min) of reached values over the cell's tile."""
    n = g.n
    Vr = V.reshape(n)
    out = np.zeros(n, np.float64)
    nt = [(v + TILE - 1) // TILE for v in n]
    for t in np.ndindex(*nt):
        sl = tuple(slice(TILE * t[k], min(TILE * t[k] + TILE, n[k])) for k in range(6))
        blk = Vr[sl].astype(np.float64)
        live = blk[blk < cap]
        out[sl] = (live.max() - live.min()) if live.size else 0.0
    return out.reshape(-1)


STORES = {"f16": (1, 2048.0), "u16": (2, 30000.0)}


@pytest.mark.parametrize("store", sorted(STORES))
def test_16bit_store_round_trip(pg, store):
    """A table loaded into a 16-bit store: its FP32 copy reads back exactly;
    the mirror holds every value within the code's resolution of its tile's
    range (binary16: 11 bits; fixed point: 16 bits, minus the quarter-span
    headroom), cap exactly, seeds exactly."""
    vstore, denom = STORES[store]
    prob = pg.configs.problem(1, shape=(12, 12, 8, 5, 5, 7))
    cap = prob.params.cap
    # a real table: a few exact sweeps
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        s.solve(20, -1.0, 0.0)
        V = s.download()
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ, vstore=vstore) as s:
        s.upload(V)
        assert np.array_equal(s.download(), V)
        W = s.download_mirror()
        assert s.kernel_info()[0] == pg.abi.PG_KERNEL_FAST
    assert np.array_equal(W >= cap, V >= cap)           # cap is out of band
    live = V < cap
    err = np.abs(W.astype(np.float64) - V)[live]
    bound = tile_ranges(prob.grid, V, cap)[live] / denom + 1e-6
    assert (err <= bound).all(), float((err - bound).max())
    assert np.all(W[prob.seeds] == 0.0)


def test_fast_sweeps_track_exact_sweeps(pg):
    """Contracted arithmetic vs the exact path, same FP32 store, 4 sweeps."""
    prob = pg.configs.problem(1)
    tabs = {}
    for kern in (pg.abi.PG_KERNEL_LDS, pg.abi.PG_KERNEL_FAST):
        with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
            s.set_kernel(kern)
            s.fill_seed(prob.seeds)
            for _ in range(4):
                s.sweep()
            tabs[kern] = s.download()
    a, b = tabs[pg.abi.PG_KERNEL_LDS].astype(np.float64), tabs[pg.abi.PG_KERNEL_FAST].astype(np.float64)
    d = np.abs(a - b)
    assert np.array_equal(a >= 60, b >= 60) or np.mean((a >= 60) != (b >= 60)) < 1e-4
    assert np.quantile(d, 0.999) < 1e-4
    assert d.mean() < 1e-5


def solve_pair(pg, shape, vstore, tol=1e-4, iters=1500, polish=3):
    """The exact FP32 table solved until no value changes, and the 16-bit
    form solved to ||dV||_inf < tol, before and after `polish` exact sweeps."""
    prob = pg.configs.problem(2, shape=shape)
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        its_e, d_e, _ = s.solve(4000, 0.0, 0.0)
        assert d_e == 0.0
        Ve = s.download()
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ, vstore=vstore) as s:
        s.fill_seed(prob.seeds)
        its_h, _, l_h = s.solve(iters, -1.0, tol)
        Vh = s.download()
        s.polish(polish)
        Vp = s.download()
        assert s.kernel_info()[0] == pg.abi.PG_KERNEL_LDS
    return prob, Ve, Vh, Vp, its_e, its_h, l_h


def error_report(name, prob, Ve, V):
    cap = prob.params.cap
    both = (Ve < cap) & (V < cap)
    e = V.astype(np.float64)[both] - Ve[both]
    a = np.abs(e)
    q = np.quantile(a, [0.5, 0.99])
    print(f"{name}: max {a.max():.2e} p50 {q[0]:.2e} p99 {q[1]:.2e} mean signed {e.mean():+.2e} "
          f"reach mismatch {np.mean((Ve < cap) != (V < cap)):.1e}")
    return a, e


def test_u16_solve_then_polish_close_to_exact(pg):
    """Precision form with the fixed-point store on a reduced problem-2 grid
    (obstacles, 125 controls), 3 exact polish sweeps, against the exact FP32
    table solved until it stops changing.

    Measured on MI355X: converges in about 150 sweeps (the exact path needs
    411 to stop changing, 280 to 1e-4); p50 2.8e-3, p99 4.8e-3, mean +2.8e-3
    after polish. The error is a bias, not noise: cells that interpolate each
    other contract by less than one code step per pass, so the rounded
    system has a band of fixed points and the iteration, coming from above,
    stops at its top. The largest differences (~0.1 s, a handful of cells)
    are discrete decisions -- probe cell, envelope edge -- that the contracted
    arithmetic takes the other way; the FP32-store fast path shows them too.
    The bounds below guard these levels."""
    prob, Ve, Vh, Vp, its_e, its_h, l_h = solve_pair(pg, (24, 24, 16, 9, 9, 11), pg.abi.PG_VSTORE_U16_TILE)
    print(f"u16: sweeps exact (to no change) {its_e}, 16-bit {its_h} (linf {l_h:.2e})")
    error_report("u16 before polish", prob, Ve, Vh)
    a, e = error_report("u16 after polish", prob, Ve, Vp)
    cap = prob.params.cap
    assert np.mean((Ve < cap) != (Vp < cap)) < 1e-4
    assert its_h < 400
    assert np.quantile(a, 0.5) < 4e-3
    assert np.quantile(a, 0.99) < 7e-3
    assert a.max() < 0.5
    assert (Vp <= Vh).all()


def test_f16_store_precision_recorded(pg):
    """The IEEE binary16 store on the same problem: 11 significant bits per
    tile leave errors a few times the fixed-point store's (recorded, not a
    target); the loose bounds only guard against a broken codec."""
    prob, Ve, Vh, Vp, its_e, its_h, l_h = solve_pair(pg, (24, 24, 16, 9, 9, 11), pg.abi.PG_VSTORE_F16_TILE,
                                                     iters=400)
    print(f"f16: 16-bit sweeps {its_h} (linf {l_h:.2e})")
    a, e = error_report("f16 after polish", prob, Ve, Vp)
    assert np.quantile(a, 0.99) < 0.05
    assert a.max() < 0.5


def test_fast_arithmetic_converged_vs_exact(pg):
    """Contracted arithmetic with an FP32 store, both solved to 1e-4: the
    arithmetic alone, without quantization. Differences come from discrete
    decisions (probe cell, envelope edge, argmin) flipping under different
    rounding -- the reference's own CPU/GPU paths differ the same way."""
    prob = pg.configs.problem(2, shape=(24, 24, 16, 9, 9, 11))
    cap = prob.params.cap
    tabs = {}
    for kern in (pg.abi.PG_KERNEL_LDS, pg.abi.PG_KERNEL_FAST):
        with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
            s.set_kernel(kern)
            s.fill_seed(prob.seeds)
            its, _, l = s.solve(1500, -1.0, 1e-4)
            tabs[kern] = (s.download(), its, l)
    Ve, Vf = tabs[pg.abi.PG_KERNEL_LDS][0], tabs[pg.abi.PG_KERNEL_FAST][0]
    both = (Ve < cap) & (Vf < cap)
    err = np.abs(Vf.astype(np.float64) - Ve)[both]
    q = np.quantile(err, [0.5, 0.9, 0.99, 0.999])
    print(f"fast-vs-exact: sweeps {tabs[pg.abi.PG_KERNEL_LDS][1]} / {tabs[pg.abi.PG_KERNEL_FAST][1]} "
          f"(linf {tabs[pg.abi.PG_KERNEL_FAST][2]:.1e}); max {err.max():.2e} quantiles {q}; "
          f"reach mismatch {np.mean((Ve < cap) != (Vf < cap)):.2e}; mean signed {np.mean((Vf.astype(np.float64) - Ve)[both]):.2e}")
    big = np.argsort(-np.abs(Vf.astype(np.float64) - Ve) * both)[:5]
    print("largest:", [(int(i), float(Ve[i]), float(Vf[i])) for i in big])
    assert err.max() < 1.0


def test_kernel_rules_for_the_16bit_store(pg):
    prob = pg.configs.problem(1, shape=(8, 8, 6, 5, 5, 5))
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ,
                      vstore=pg.abi.PG_VSTORE_U16_TILE) as s:
        s.fill_seed(prob.seeds)
        with pytest.raises(pg.abi.PeregrineError):
            s.set_kernel(pg.abi.PG_KERNEL_LDS)
        s.sweep()
        s.polish(1)
        s.set_kernel(pg.abi.PG_KERNEL_GLOBAL)
        s.sweep()
