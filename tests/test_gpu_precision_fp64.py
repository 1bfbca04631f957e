"""The precision form's acceptance check against FP64, on a reduced
problem-2 grid (24x24x16x9x9x11: obstacles, 125 controls, the plain scheme).

  1. The FP64 truth: the device's diagonal-solved FP64 Jacobi table, solved
     to ||dV||inf < 1e-10 s, certified by the CPU FP64 oracle: one oracle
     backup of 6,000 seeded reached states reproduces it to 1e-9 s.
  2. The exact FP32 table (the reference's arithmetic) solved until a sweep
     changes nothing: where it is more than 5e-3 s from FP64 the FP32
     arithmetic itself takes a discrete decision (a probe cell, an envelope
     edge) the other way, which no faster form can be asked to undo. Those
     states are excluded, counted and listed (measured: 391 of 8.2M).
  3. The precision form: 16-bit sweeps to its own ||dV||inf < 1e-3 s, then
     exact FP32 polish sweeps until one moves no state by 1e-3 s or more --
     the metric's stop.

Against the FP64 table, over the states both reach and outside the excluded
ones, the precision form's maximum error is at most 5e-3 s (measured
4.7e-3; p99 9.7e-4). For comparison the exact FP32 table stopped by the same
rule is reported too: it is not yet converged there (max 1.5e-2, p99 5.0e-3).
Argmin agreement with the FP64 policy is reported and bounded.
"""
import numpy as np
import pytest

from conftest import nthreads, oracle_problem

pytestmark = pytest.mark.gpu

SHAPE = (24, 24, 16, 9, 9, 11)
TOL = 1e-3
BAR = 5e-3


def polished_precision_form(pg, prob, max_polish=60):
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ,
                      vstore=pg.abi.PG_VSTORE_U16_TILE) as s:
        s.fill_seed(prob.seeds)
        its, _, l16 = s.solve(2000, -1.0, TOL)
        s.polish(0)
        trace = []
        for k in range(max_polish):
            r = s.sweep(k, k % 2 == 1)
            trace.append(r.linf)
            if r.linf < TOL:
                break
        return s.download(), s.policy(), its, l16, trace


def test_precision_form_within_5ms_of_fp64(pg, orc):
    prob = pg.configs.problem(2, shape=SHAPE)
    cap = prob.params.cap
    n = prob.grid.ncells
    with pg.HipSolver(prob.grid, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        its32, _, l32 = s.solve(3000, -1.0, TOL)
        Ve = s.download().astype(np.float64)
        its_c, d_c, _ = s.solve(8000, 0.0, 0.0)
        assert d_c == 0.0
        Vc = s.download().astype(np.float64)
        its64, l64 = s.solve_f64(5000, 1e-10)
        V64 = s.download_f64()
        P64 = s.policy_f64()
        r64, _ = s.residual_f64(0)
    print(f"FP64: {its64} diagonal sweeps, linf {l64:.1e}, device residual {r64:.1e}; "
          f"FP32: {its32} sweeps to linf {l32:.1e}, {its_c} more until nothing changes", flush=True)
    assert l64 < 1e-10 and r64 < 1e-8

    # certify the FP64 truth with the CPU FP64 oracle's own backups
    op = oracle_problem(orc, prob)
    rng = np.random.default_rng(3)
    seeds = np.zeros(n, bool)
    seeds[prob.seeds] = True
    reached = np.nonzero((V64 < cap) & ~seeds)[0]
    cells = np.sort(rng.choice(reached, 6000, replace=False))
    TV, _ = op.backup_cells_f64(V64, cells, nthreads=nthreads())
    cert = float(np.max(np.abs(TV - V64[cells])))
    print(f"oracle FP64 backups of 6000 reached states: max |T V - V| {cert:.1e}", flush=True)
    assert cert < 1e-9

    Vp, Pp, its16, l16, trace = polished_precision_form(pg, prob)
    print(f"precision form: {its16} 16-bit sweeps (linf {l16:.1e}), {len(trace)} polish sweeps, "
          f"exact linf {['%.1e' % v for v in trace]}", flush=True)
    Vp = Vp.astype(np.float64)
    both = (Vp < cap) & (V64 < cap) & (Vc < cap)
    ec = np.abs(Vc - V64)
    ee = np.abs(Ve - V64)
    ep = np.abs(Vp - V64)
    fp32_off = both & (ec > BAR)
    keep = both & ~fp32_off
    q = np.quantile(ep[both], [0.5, 0.99, 0.999])
    qe = np.quantile(ee[both], [0.5, 0.99, 0.999])
    print(f"exact FP32 at the 1e-3 stop vs FP64: max {ee[both].max():.2e} (outside the excluded states "
          f"{ee[keep].max():.2e}) p50 {qe[0]:.1e} p99 {qe[1]:.1e} p99.9 {qe[2]:.1e}", flush=True)
    print(f"precision form vs FP64: max {ep[both].max():.2e} p50 {q[0]:.1e} p99 {q[1]:.1e} p99.9 {q[2]:.1e}; "
          f"max outside the excluded states: {ep[keep].max():.2e}", flush=True)
    off = np.nonzero(fp32_off)[0]
    print(f"{off.size} states where the converged exact FP32 table is more than {BAR:g} s from FP64 "
          f"(cell, FP32 - FP64, precision - FP64): "
          f"{[(int(i), round(float(Vc[i] - V64[i]), 4), round(float(Vp[i] - V64[i]), 4)) for i in off[:12]]}",
          flush=True)
    # argmin agreement over free, reached, non-target states (near-ties,
    # where either choice is optimal to rounding, count as disagreements)
    blocked = np.repeat(np.asarray(prob.occ, bool), n // prob.occ.size)
    live = (V64 < cap) & ~seeds & ~blocked
    agree = float(np.mean(Pp[live] == P64[live]))
    print(f"argmin agreement with the FP64 policy: {agree:.5f} of {int(live.sum())} states", flush=True)
    assert np.mean((Vp < cap) != (V64 < cap)) < 1e-4
    assert off.size <= 1e-4 * both.sum()
    assert ep[keep].max() <= BAR
    assert agree > 0.99
