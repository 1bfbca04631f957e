"""The reference's whole-solve self-tests, run through the HIP path.

Both compare a solve with a closed form that nothing in the solver knows,
so together they check the chain -- control set, integration, probes,
interpolation, seeding, the iteration and its stopping rule:

  * test 1, minimum time (solver/src/SelfTest.jl:90-199): a double
    integrator a = 100 u from rest to rest over d has the bang-bang time
    2 sqrt(d/a). Three schemes -- the plain fixed-dt scan, the horizon
    ladder with RK2 and adaptive probes, and that plus the rotating slice
    and pattern search -- with the reference's thresholds: excess over the
    closed form split into a fixed part (< 0.45 s) and a per-second part
    (< 20%), the ladder no worse than fixed dt by 2%, the fast search no
    worse than the ladder by 8%;
  * test 16, the escape pass (SelfTest.jl:1372-1510): behind an obstacle
    slab, the escape from rest at x = 25 cm must aim at the far edge
    (sqrt(2 d/a) over 15 cm), not the near one (5 cm); cells with a route
    keep their values bit for bit; cells walled off get no escape.

The values are read with the Kuhn interpolant the reference reads them with
(`interp_kuhn`), evaluated by the oracle on the downloaded table.
"""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

AMAX = 100.0
CAP = 30.0
N = (81, 5, 4, 33, 5, 5)


def double_integrator(pg):
    Z = (0.0,) * 9
    return pg.Model((AMAX, 0, 0, 0, AMAX, 0, 0, 0, AMAX), Z, (0, 0, 0), Z, Z, (0, 0, 0), (5, 5, 0.15), 0.0)


def grid(pg):
    return pg.Grid6.build(N, (0, 0, -math.pi, -100, -100, -2), (200, 200, math.pi, 100, 100, 2))


def seeds(pg, g):
    ttol = [float(np.float32(g.step[k]) * np.float32(0.5)) for k in range(6)]
    return pg.target_cells(g, (100, 100, 0, 0, 0, 0), ttol)


def kuhn_at(orc, g, V, st):
    st = np.asarray(st, np.float32)
    grad = np.zeros(6, np.float32)
    return float(orc.lib().orc_interp_kuhn_f32(V.ctypes.data, ctypes.byref(g.c()), st.ctypes.data, CAP, 0,
                                               grad.ctypes.data))


def solved(pg, g, m, occ, p, warm, tidx):
    ctl = pg.control_set(4)
    with pg.HipSolver(g, m, p, ctl, occ) as s:
        if warm:
            s.set_warm_start(True)
        it, last = pg.solve_value(s, tidx, iters=600, tol=1e-4)
        print(f"  solve: {it} sweeps, last summed decrease {last:.3g}", flush=True)
        return s.download()


def test_min_time_closed_form(pg, orc):
    g = grid(pg)
    m = double_integrator(pg)
    occ = np.zeros(N[0] * N[1] * N[2], np.uint8)
    tidx = seeds(pg, g)
    common = dict(dt=0.05, nsub=4, cap=CAP, tau_ratio=2.0, tau_min=0.004, tau_max=0.5, delta0=0.35,
                  nearest=False, vclamp=False, simplex=True)
    legacy = pg.Params(**common, checks=3, adaptive_checks=False, ntau=1, cfl=0.0, hmax=0.0125, ncoarse=0,
                       rounds=0, rk2=False)
    ladder = pg.Params(**common, checks=2, adaptive_checks=True, ntau=4, cfl=1.0, hmax=0.0125, ncoarse=0,
                       rounds=0, rk2=True)
    now = pg.Params(**common, checks=2, adaptive_checks=True, ntau=4, cfl=1.0, hmax=0.0125, ncoarse=8,
                    rounds=2, rk2=True)
    ds = (20.0, 40.0, 60.0)
    want = np.array([2.0 * math.sqrt(d / AMAX) for d in ds])
    got = {}
    for name, p in (("fixed-dt", legacy), ("ladder", ladder), ("fast-search", now)):
        print(f"min time, {name} scheme", flush=True)
        V = solved(pg, g, m, occ, p, p.ncoarse > 0, tidx)
        got[name] = np.array([kuhn_at(orc, g, V, (100.0 - d, 100.0, 0, 0, 0, 0)) for d in ds])
        print(f"  d = {ds}: exact {np.round(want, 3)}, solved {np.round(got[name], 3)}", flush=True)
    ml, md, mn = (got[k].mean() for k in ("fixed-dt", "ladder", "fast-search"))
    assert mn <= md * 1.08, (mn, md)
    assert md <= ml * 1.02, (md, ml)
    ex = got["fast-search"] - want
    slope = (ex[2] - ex[0]) / (want[2] - want[0])
    intercept = ex[0] - slope * want[0]
    print(f"  excess = {intercept:.3f} s fixed + {100 * slope:.1f}% of the journey", flush=True)
    assert intercept < 0.45
    assert slope < 0.20
    for name in got:                              # every scheme lands near the closed form
        assert np.all(got[name] > 0.8 * want) and np.all(got[name] < 1.6 * want), (name, got[name])


def test_escape_closed_form(pg):
    g = grid(pg)
    m = double_integrator(pg)
    nxy = N[0] * N[1] * N[2]
    xs = g.axisvalue(0, np.arange(N[0])).astype(np.float64)
    slab = [i for i in range(N[0]) if 20.0 <= xs[i] <= 40.0]
    assert len(slab) >= 3
    occ = np.zeros((N[0], N[1] * N[2]), np.uint8)
    occ[slab] = 1
    occ = occ.reshape(nxy)
    tidx = seeds(pg, g)
    p = pg.Params(dt=0.05, nsub=4, cap=CAP, checks=2, adaptive_checks=True, ntau=4, cfl=1.0, hmax=0.0125,
                  ncoarse=0, rounds=2, rk2=True, tau_ratio=2.0, tau_min=0.004, tau_max=0.5, delta0=0.35,
                  nearest=False, vclamp=False, simplex=True)
    with pg.HipSolver(g, m, p, pg.control_set(4), occ) as s:
        it, _ = pg.solve_value(s, tidx, iters=600, tol=1e-4)
        before = s.download()
        eit, _, _ = s.solve_escape(200, tol=1e-4)
        after = s.download()
    print(f"escape: value {it} sweeps, escape {eit} sweeps", flush=True)
    term = (before >= 0) & (before < CAP)
    assert term.any()
    assert np.array_equal(after[term].view(np.uint32), before[term].view(np.uint32))
    esc = after < 0
    assert esc.any()
    assert np.all(np.isfinite(after[esc])) and np.all(-after[esc] > 0) and np.all(-after[esc] <= CAP + 1e-3)

    i = int(np.argmin(np.abs(xs - 25.0)))
    j = int(round((100.0 - float(g.lo[1])) / float(g.step[1])))
    i4, i5, i6 = (N[3] - 1) // 2, (N[4] - 1) // 2, (N[5] - 1) // 2
    v = float(after[int(g.flatten(i, j, 0, i4, i5, i6))])
    assert v < 0
    got = -v
    want_far = math.sqrt(2 * 15.0 / AMAX)
    want_near = math.sqrt(2 * 5.0 / AMAX)
    print(f"  x = 25 cm: far edge {want_far:.3f} s, near edge {want_near:.3f} s, solved {got:.3f} s", flush=True)
    assert got > (want_far + want_near) / 2
    assert got <= want_far * 1.6

    blk = N[3] * N[4] * N[5]
    low = after.reshape(N[0], -1)[: slab[0]]
    assert not (low < 0).any(), f"{int((low < 0).sum())} cells behind the slab escaped through it"
    assert low.size == slab[0] * N[1] * N[2] * blk
