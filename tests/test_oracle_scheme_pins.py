"""Pins of the oracle's plain-scheme options: nearest-node lookups,
adaptive probe counts and the rotating slice of the control lattice
(PeregrineSolver.jl:424-428, 1127-1133, 1263-1275). Each reduces to a
brute-force statement of what the option means."""
import ctypes
import dataclasses
import math

import numpy as np
import pytest

from conftest import oracle_problem


def test_nearest_reads_the_rounded_node(pg, orc):
    g = pg.Grid6.build((5, 6, 8, 4, 5, 3), (0, 0, -math.pi, -100, -100, -4), (80, 100, math.pi, 100, 100, 4))
    rng = np.random.default_rng(5)
    V = rng.uniform(0, 10, g.ncells).astype(np.float32)
    gc = g.c()
    L = orc.lib()
    n = np.array(g.n)
    lo = np.array(g.lo, np.float32)
    step = np.array(g.step, np.float32)
    for _ in range(300):
        f = rng.uniform(0, 1, 6) * (n - 1)
        f[2] = rng.uniform(0, n[2])
        st = (lo + f.astype(np.float32) * step).astype(np.float32)
        got = L.orc_interp_f32(V.ctypes.data, ctypes.byref(gc), st.ctypes.data, 1, 30.0, 0)
        fc = [(np.float32(st[k]) - lo[k]) / step[k] for k in range(6)]
        fc[2] = fc[2] - np.floor(fc[2] / np.float32(n[2])) * np.float32(n[2])
        i = [int(np.rint(fc[k])) for k in range(6)]
        i[2] %= n[2]
        idx = 0
        for k in range(6):
            idx = idx * n[k] + i[k]
        assert got == V[idx]


def test_adaptive_probes_catch_a_thin_wall(pg, orc):
    """A 60 cm step over a one-cell wall: the one fixed probe (at the
    successor) misses it, the adaptive count (a probe every half cell)
    rejects the move."""
    g = pg.Grid6.build((11, 5, 4, 9, 3, 3), (0, 0, -math.pi, -2000, -100, -4), (100, 40, math.pi, 2000, 100, 4))
    occ = np.zeros(11 * 5 * 4, np.uint8)
    occ.reshape(11, 5, 4)[5, :, :] = 1
    V = np.ones(g.ncells, np.float32)
    m = pg.Model((100, 0, 0, 0, 100, 0, 0, 0, 10), (0.0,) * 9, (0, 0, 0), (0.0,) * 9, (0.0,) * 9, (0, 0, 0),
                 (5, 5, 0.15), 0.0)
    mc = m.c_struct()
    L = orc.lib()
    s = np.array([20.0, 20.0, 0.0, 1200.0, 0.0, 0.0], np.float32)
    u = np.zeros(3, np.float32)
    base = pg.Params.plain(dt=0.05, nsub=4, checks=1)
    out = {}
    for adaptive in (False, True):
        p = dataclasses.replace(base, adaptive_checks=adaptive).c()
        out[adaptive] = L.orc_lookahead_f32(V.ctypes.data, occ.ctypes.data, ctypes.byref(g.c()), ctypes.byref(mc),
                                            s.ctypes.data, u.ctypes.data, 0.05, 4, ctypes.byref(p))
    assert out[False] < base.cap
    assert out[True] == np.float32(base.cap)


def test_rotating_slice_covers_the_lattice_in_a_turn(pg, orc):
    """With a slice of 5 of problem 1's 26 moving controls, the best value
    over one turn of phases (ceil(26/5) = 6) is the full scan's."""
    prob = pg.configs.problem(1, shape=(8, 8, 6, 5, 5, 5))
    sliced = dataclasses.replace(prob, params=dataclasses.replace(prob.params, ncoarse=5))
    rng = np.random.default_rng(9)
    V = rng.uniform(0, 5, prob.grid.ncells).astype(np.float32)
    full = oracle_problem(orc, prob)
    part = oracle_problem(orc, sliced)
    for idx in rng.choice(prob.grid.ncells, 60, replace=False):
        want, _ = full.cell_update(V, int(idx))
        got = [part.cell_update(V, int(idx), phase=ph)[0] for ph in range(6)]
        assert min(got) == want
        assert all(v >= want for v in got)


def test_cfl_horizon_takes_the_tightest_term(pg, orc):
    """cfl_tau: tau_max, cfl * step / |B| per velocity axis, cfl * min xy step
    / speed and cfl * heading step / |w|, the smallest, clamped to
    [tau_min, tau_max]. On a grid with 10 cm x/y steps, 50 cm/s velocity
    steps and a 0.5 rad heading step, B = diag(1000, 1000, 100), cfl 0.5:
    the acceleration terms give 0.025 s, a 400 cm/s speed 0.0125 s, a 10 rad/s
    spin 0.5 * 2pi/13 / 10 = 0.0242 s."""
    L = orc.lib()
    g = pg.Grid6.build((11, 11, 13, 9, 9, 9), (0, 0, -math.pi, -200, -200, -20), (100, 100, math.pi, 200, 200, 20))
    assert g.step[0] == pytest.approx(10) and g.step[3] == pytest.approx(50)
    m = pg.Model((1000, 0, 0, 0, 1000, 0, 0, 0, 100), (0.0,) * 9, (0, 0, 0), (0.0,) * 9, (0.0,) * 9, (0, 0, 0),
                 (5, 5, 0.15), 0.0).c_struct()
    p = dataclasses.replace(pg.Params.plain(), cfl=0.5, tau_min=0.004, tau_max=0.5)
    pc = p.c()
    gc = g.c()
    tau = lambda vx, vy, w: L.orc_cfl_tau(ctypes.byref(gc), ctypes.byref(m), vx, vy, w, ctypes.byref(pc))
    h_step = float(np.float32(g.step[2]))
    assert tau(0.0, 0.0, 0.0) == pytest.approx(0.025, rel=1e-6)             # 0.5 * 50 / 1000
    assert tau(400.0, 0.0, 0.0) == pytest.approx(0.0125, rel=1e-6)          # 0.5 * 10 / 400
    assert tau(0.0, 0.0, 10.0) == pytest.approx(0.5 * h_step / 10.0, rel=1e-6)
    assert tau(0.0, 0.0, 100.0) == pytest.approx(0.004)                   # below tau_min
    slow = dataclasses.replace(p, tau_max=0.01).c()
    assert L.orc_cfl_tau(ctypes.byref(gc), ctypes.byref(m), 0.0, 0.0, 0.0, ctypes.byref(slow)) == pytest.approx(0.01)
    floor = dataclasses.replace(p, tau_min=0.02).c()
    assert L.orc_cfl_tau(ctypes.byref(gc), ctypes.byref(m), 400.0, 0.0, 0.0, ctypes.byref(floor)) == pytest.approx(0.02)
