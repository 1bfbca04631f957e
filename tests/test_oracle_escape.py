"""The CPU restatement of the escape pass (gpu_escape_kernel!, Gpu.jl:51-67;
EscapeView / OccView, PeregrineSolver.jl:756-856), pinned on the CPU.

The escape backup of a free cell is the plain backup of a transformed table
(obstacle cells at cap, escape times as themselves, real routes as 0); a
blocked cell's is the plain backup of the unmasked transform with the
occupancy waived. The plain backup is pinned elsewhere, so these reductions
pin the escape reading. A short solve then checks the pass's invariants.
"""
import dataclasses

import numpy as np
import pytest

from conftest import nthreads, oracle_problem


def small_problem(pg):
    """Problem 2's six obstacles on a 10x10x8x5x5x5 grid, 27 controls."""
    prob = pg.configs.problem(2, shape=(10, 10, 8, 5, 5, 5))
    return dataclasses.replace(prob, ctl=pg.configs.cube_controls(pg.configs.LEVELS3))


def transform(V, occ_cells, cap, mask):
    out = np.where(V < 0, -V, np.where(V < cap, 0.0, cap)).astype(np.float32)
    if mask:
        out[occ_cells] = cap
    return out


def test_escape_backup_is_plain_backup_of_transformed_table(pg, orc):
    prob = small_problem(pg)
    g = prob.grid
    cap = np.float32(prob.params.cap)
    rng = np.random.default_rng(11)
    n = g.ncells
    # a mix of routes, escape times and unsolved cells
    V = rng.uniform(0.0, 3.0, n).astype(np.float32)
    V[rng.random(n) < 0.3] = cap
    esc = rng.random(n) < 0.3
    V[esc] = -rng.uniform(0.1, 5.0, esc.sum()).astype(np.float32)
    blk = g.n[3] * g.n[4] * g.n[5]
    occ = np.asarray(prob.occ, np.uint8)
    occ_cells = np.repeat(occ != 0, blk)
    assert occ_cells.any() and (~occ_cells).any()
    op = oracle_problem(orc, prob)
    free_op = oracle_problem(orc, dataclasses.replace(prob, occ=np.zeros_like(occ)))
    masked = transform(V, occ_cells, cap, True)
    unmasked = transform(V, occ_cells, cap, False)
    free = np.nonzero(~occ_cells)[0]
    blocked = np.nonzero(occ_cells)[0]
    for idx in rng.choice(free, 200, replace=False):
        assert op.cell_update(V, int(idx), esc=1) == op.cell_update(masked, int(idx), esc=0)
    for idx in rng.choice(blocked, 100, replace=False):
        assert op.cell_update(V, int(idx), esc=1) == free_op.cell_update(unmasked, int(idx), esc=0)


def test_escape_pass_invariants(pg, orc):
    """After 60 value sweeps: routes are never written, every escape time is
    in [dt, cap) and only falls, and blocked cells find a way out."""
    prob = small_problem(pg)
    op = oracle_problem(orc, prob)
    cap = np.float32(prob.params.cap)
    V = np.full(prob.grid.ncells, cap, np.float32)
    V[prob.seeds] = 0.0
    W = np.empty_like(V)
    for _ in range(60):
        op.sweep(V, W, nthreads=nthreads())
        W[prob.seeds] = 0.0
        V, W = W, V
    routes = (V >= 0) & (V < cap)
    E = V.copy()
    prev_esc = np.where(E < 0, -E, cap)
    for it in range(25):
        op.sweep(E, W, nthreads=nthreads(), esc=1)
        E, W = W, E
        assert np.array_equal(E[routes], V[routes])
        e = np.where(E < 0, -E, cap)
        assert np.all(e <= prev_esc)
        prev_esc = e
    got = E[E < 0]
    assert got.size > 0 and np.all(-got >= prob.params.dt) and np.all(-got < cap)
    blk = prob.grid.n[3] * prob.grid.n[4] * prob.grid.n[5]
    occ_cells = np.repeat(np.asarray(prob.occ, np.uint8) != 0, blk)
    assert np.count_nonzero(E[occ_cells] < 0) > 0
