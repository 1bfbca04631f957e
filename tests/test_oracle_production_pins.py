"""Pins of the oracle's production-scheme steps (PeregrineSolver.jl:
1240-1343, 728-744, 1165-1170): the traction gain's tanh, the horizon
ladder, the pattern search and the warm start. Each is checked against an
explicit enumeration of the candidates the step means, evaluated one by one
through the oracle's single-command lookahead, at cells whose velocity node
is zero (so the body-frame velocity is exactly zero and the start state is
known without the oracle's sincos)."""
import ctypes
import dataclasses
import math

import numpy as np
import pytest

from conftest import oracle_problem

F = np.float32


def test_tanh_is_correctly_rounded(orc):
    """The oracle's float tanh equals the C library's double tanh rounded to
    float on a dense sample of [1e-6, 30] and at the ends of its branches."""
    L = orc.lib()
    rng = np.random.default_rng(11)
    xs = np.concatenate([rng.uniform(0, 1, 100000), rng.uniform(1, 30, 50000),
                         np.exp(rng.uniform(math.log(1e-6), 0, 50000)),
                         [1e-6, 0.25, 0.2499999, 0.2500001, 9.01, 20.0, 29.9]]).astype(F)
    got = np.array([L.orc_tanh_f32(float(x)) for x in xs], F)
    want = np.array([F(math.tanh(float(x))) for x in xs], F)
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} of {xs.size} differ"


@pytest.fixture(scope="module")
def case(pg):
    """Problem 2's model (knee off) and obstacles on a small grid whose
    velocity axes have a zero node, a random table, and free cells at zero
    velocity."""
    prob = pg.configs.problem(2, shape=(12, 12, 8, 5, 5, 7))
    g = prob.grid
    rng = np.random.default_rng(3)
    V = rng.uniform(0.5, 8.0, g.ncells).astype(F)
    z = [int(np.nonzero(g.axisvalue(k, np.arange(g.n[k])) == 0)[0][0]) for k in (3, 4)]
    occ = np.asarray(prob.occ, np.uint8).reshape(g.n[0], g.n[1], g.n[2])
    cells = []
    while len(cells) < 40:
        i = [int(rng.integers(0, g.n[k])) for k in range(6)]
        i[3], i[4] = z
        if occ[i[0], i[1], i[2]]:
            continue
        cells.append(tuple(i))
    return prob, V, cells


def start_state(g, i):
    """The cell's start state: axis values, body velocity (0, 0)."""
    return np.array([g.axisvalue(0, i[0]), g.axisvalue(1, i[1]), g.axisvalue(2, i[2]), 0.0, 0.0,
                     g.axisvalue(5, i[5])], F)


def with_params(prob, **kw):
    return dataclasses.replace(prob, params=dataclasses.replace(prob.params, **kw))


def evaluator(orc, prob, V, p):
    L = orc.lib()
    gc, mc, pc = prob.grid.c(), prob.model.c_struct(), p.c()
    occ = np.ascontiguousarray(prob.occ, np.uint8)

    def look(s, u, tau):
        u = np.asarray(u, F)
        ns = L.orc_substeps_for(F(tau), ctypes.byref(pc))
        return F(L.orc_lookahead_f32(V.ctypes.data, occ.ctypes.data, ctypes.byref(gc), ctypes.byref(mc),
                                     s.ctypes.data, u.ctypes.data, F(tau), ns, ctypes.byref(pc)))

    def tau0(i):
        if p.cfl <= 0:
            return F(p.dt)
        return F(L.orc_cfl_tau(ctypes.byref(gc), ctypes.byref(mc), 0.0, 0.0, float(prob.grid.axisvalue(5, i[5])),
                               ctypes.byref(pc)))
    return look, tau0


def base_best(orc, prob, V, idx):
    """The lattice scan alone: value and winning command."""
    plain = with_params(prob, rounds=0, ntau=1)
    v, k = oracle_problem(orc, plain).cell_update(V, idx)
    return F(v), np.asarray(prob.ctl[k], F)


def test_ladder_is_the_best_command_at_each_rung(pg, orc, case):
    """ntau rungs tau0 * 0.5 * ratio^(k-1), rung 2 skipped, each clamped to
    [tau_min, tau_max], plus the dt rung when cfl > 0, all with the lattice
    winner: the value is the minimum of the lattice value and the rungs."""
    prob, V, cells = case
    for kw in (dict(ntau=4, tau_ratio=2.0), dict(ntau=5, tau_ratio=1.7, cfl=0.8, tau_min=0.01, tau_max=0.3)):
        pp = with_params(prob, **kw)
        op = oracle_problem(orc, pp)
        look, tau0 = evaluator(orc, pp, V, pp.params)
        for i in cells:
            idx = int(pp.grid.flatten(*i))
            best, b = base_best(orc, with_params(pp), V, idx)
            s = start_state(pp.grid, i)
            want = best
            t = F(tau0(i)) * F(0.5)
            for k in range(1, pp.params.ntau + 1):
                if k != 2:
                    tc = min(max(t, F(pp.params.tau_min)), F(pp.params.tau_max))
                    want = min(want, look(s, b, tc))
                t = F(t * F(pp.params.tau_ratio))
            if pp.params.cfl > 0:
                tc = min(max(F(pp.params.dt), F(pp.params.tau_min)), F(pp.params.tau_max))
                want = min(want, look(s, b, tc))
            got, _ = op.cell_update(V, idx)
            assert F(got) == want, (i, got, want)


def proj_oct(c):
    s = F(F(abs(c[0]) + abs(c[1])) + abs(c[2]))
    if s < F(1e-6):
        return np.zeros(3, F)
    r = F(F(1) / s)
    return np.array([c[0] * r, c[1] * r, c[2] * r], F)


def test_pattern_search_walks_the_octahedron(pg, orc, case):
    """rounds of six moves (+d, -d on fwd, strafe, turn in that order) from
    the best command so far, each projected onto |u|_1 = 1 and skipped when
    zero, d halving per round; a strictly better move is taken at once."""
    prob, V, cells = case
    pp = with_params(prob, rounds=3, delta0=0.35)
    op = oracle_problem(orc, pp)
    look, tau0 = evaluator(orc, pp, V, pp.params)
    moved = 0
    for i in cells:
        idx = int(pp.grid.flatten(*i))
        best, b = base_best(orc, pp, V, idx)
        s = start_state(pp.grid, i)
        t0 = tau0(i)
        d = F(pp.params.delta0)
        for _ in range(pp.params.rounds):
            for a in range(3):
                for sg in (d, -d):
                    c = b.copy()
                    c[a] = F(c[a] + sg)
                    c = proj_oct(c)
                    if F(F(abs(c[0]) + abs(c[1])) + abs(c[2])) < F(1e-6):
                        continue
                    v = look(s, c, t0)
                    if v < best:
                        best, b = v, c
                        moved += 1
            d = F(d * F(0.5))
        got, _ = op.cell_update(V, idx)
        assert F(got) == best, (i, got, best)
    assert moved > 0


def test_warm_start_store(pg, orc, case):
    """A zero store changes nothing; a sweep stores the winner of every cell
    it improves, quantised to 1/127; a stored command is a candidate."""
    prob, V, cells = case
    op = oracle_problem(orc, prob)
    n = prob.grid.ncells
    plain = np.empty_like(V)
    op.sweep(V, plain)
    wpol = np.zeros(3 * n, np.int8)
    warm = np.empty_like(V)
    op.sweep_warm(V, warm, wpol)
    assert np.array_equal(warm, plain)
    pol = np.empty(n, np.uint8)
    op.sweep(V, np.empty_like(V), pol=pol)
    improved = np.nonzero(plain < V)[0]
    q = np.rint(np.clip(prob.ctl[pol[improved]], -1, 1) * F(127)).astype(np.int8)
    assert np.array_equal(wpol.reshape(-1, 3)[improved], q)
    assert not wpol.reshape(-1, 3)[plain >= V].any()
    # a stored command is evaluated after coast, before the lattice
    look, tau0 = evaluator(orc, prob, V, prob.params)
    rng = np.random.default_rng(5)
    store = rng.integers(-127, 128, size=3 * n).astype(np.int8)
    out = np.empty_like(V)
    op.sweep_warm(V, out, store.copy())
    for i in cells:
        idx = int(prob.grid.flatten(*i))
        qv = store[3 * idx:3 * idx + 3].astype(F) * F(F(1) / F(127))
        best, _ = base_best(orc, prob, V, idx)
        want = min(best, look(start_state(prob.grid, i), qv, tau0(i)))
        assert out[idx] == min(want, V[idx]), i
