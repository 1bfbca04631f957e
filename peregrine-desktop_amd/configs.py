"""The five seeded benchmark problems of this solver.

Every problem is a synthetic instance of the reference's value solve: a 6-D
grid with the reference's axis conventions, a drivetrain model drawn from a
fixed seeded distribution, cube control sets with coast first, axis-aligned
rectangular obstacles rasterised for a point robot, and a disc target at the
field centre. Seeds: model 1, obstacles 2, error subsample 3, adjoint starts
4; problem c draws from numpy PCG64(seed + 100*c). The precision problem
(3) is problem 2 stored in FP16, so it reuses problem 2's draws.

  1  16x16x12x5x5x7      537,600 states, 27 controls, no obstacles
  2  64x64x48x13x13x17   564,854,784 states, 125 controls, 6 obstacles
  3  = 2, V stored FP16 with per-tile scale/offset
  4  96x96x64x17x17x21   3,579,641,856 states, 125 controls, out of core
  5  32x32x24x9x9x11     21,897,216 states, 27 controls, adjoint
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .model import Grid6, Model, Params, cube_controls, disc_target_cells
from .occupancy import build_occupancy, rect

FIELD = 365.76
VMAX = 180.0
WMAX = 6.0
TARGET = (182.88, 182.88, 0.0, 0.0, 0.0, 0.0)
TARGET_RADIUS = 10.0
TARGET_DH = 0.1

SHAPES = {
    1: (16, 16, 12, 5, 5, 7),
    2: (64, 64, 48, 13, 13, 17),
    3: (64, 64, 48, 13, 13, 17),
    4: (96, 96, 64, 17, 17, 21),
    5: (32, 32, 24, 9, 9, 11),
}
LEVELS3 = (-1.0, 0.0, 1.0)
LEVELS5 = (-1.0, -0.5, 0.0, 0.5, 1.0)
NOBST = {1: 0, 2: 6, 3: 6, 4: 6, 5: 0}


def rng_for(seed: int, cfg: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed + 100 * cfg))


def seed_cfg(cfg: int) -> int:
    """The draw index: the precision problem shares problem 2's draws."""
    return 2 if cfg == 3 else cfg


def random_model(cfg: int) -> Model:
    """B = diag(U[150,250], U[150,250], U[8,14]) with N(0, (0.05 B_ii)^2)
    off-diagonals in row i; A = -diag(U[2,5]); q = 0;
    S = -diag(U[0,10], U[0,10], U[0,1]); D = -diag(U[0,5e-3], U[0,5e-3],
    U[0,5e-2]); c ~ N(0, diag(0.5, 0.5, 0.02)^2); eps (5, 5, 0.15); no knee.
    Draw order: B diagonal, B off-diagonals row-major, A, S, D, c."""
    r = rng_for(1, seed_cfg(cfg))
    bd = np.array([r.uniform(150, 250), r.uniform(150, 250), r.uniform(8, 14)])
    B = np.diag(bd)
    for i in range(3):
        for j in range(3):
            if i != j:
                B[i, j] = r.normal(0.0, 0.05 * bd[i])
    A = -np.diag([r.uniform(2, 5) for _ in range(3)])
    S = -np.diag([r.uniform(0, 10), r.uniform(0, 10), r.uniform(0, 1)])
    D = -np.diag([r.uniform(0, 5e-3), r.uniform(0, 5e-3), r.uniform(0, 5e-2)])
    c = np.array([r.normal(0, 0.5), r.normal(0, 0.5), r.normal(0, 0.02)])
    return Model(tuple(B.ravel()), tuple(A.ravel()), (0.0, 0.0, 0.0), tuple(S.ravel()),
                 tuple(D.ravel()), tuple(c), (5.0, 5.0, 0.15), 0.0)


def random_obstacles(cfg: int, n: int):
    """n axis-aligned rectangles, sides U[20,60] cm, centres uniform in the
    field, redrawn while the rectangle comes within the target disc."""
    r = rng_for(2, seed_cfg(cfg))
    out = []
    while len(out) < n:
        w, h = r.uniform(20, 60), r.uniform(20, 60)
        cx, cy = r.uniform(0, FIELD), r.uniform(0, FIELD)
        dx = max(abs(TARGET[0] - cx) - w / 2, 0.0)
        dy = max(abs(TARGET[1] - cy) - h / 2, 0.0)
        if math.hypot(dx, dy) <= TARGET_RADIUS:
            continue
        out.append((cx, cy, w, h))
    return out


@dataclass
class Problem:
    cfg: int
    grid: Grid6
    model: Model
    params: Params
    ctl: np.ndarray
    rects: list
    occ: np.ndarray = field(repr=False)
    seeds: np.ndarray = field(repr=False)

    @property
    def nctl(self) -> int:
        return int(self.ctl.shape[0])


def make_grid(shape) -> Grid6:
    return Grid6.build(shape, (0.0, 0.0, -math.pi, -VMAX, -VMAX, -WMAX),
                       (FIELD, FIELD, math.pi, VMAX, VMAX, WMAX))


def problem(cfg: int, shape=None, with_occ: bool = True) -> Problem:
    """Problem `cfg`; `shape` overrides the grid (same model, obstacles and
    controls) for scaled-down parity cases."""
    shape = SHAPES[cfg] if shape is None else tuple(shape)
    g = make_grid(shape)
    m = random_model(cfg)
    ctl = cube_controls(LEVELS5 if cfg in (2, 3, 4) else LEVELS3)
    rects = random_obstacles(cfg, NOBST[cfg])
    if with_occ:
        occ = build_occupancy(g, [rect(*r) for r in rects], 0.0, None, 3, 0.0)
    else:
        occ = np.zeros(g.n[0] * g.n[1] * g.n[2], dtype=np.uint8)
    seeds = disc_target_cells(g, TARGET, TARGET_RADIUS, TARGET_DH)
    return Problem(cfg, g, m, Params.plain(), ctl, rects, occ, seeds)


def algorithmic_bytes_per_update(nctl: int, bv: int) -> int:
    """Gathered bytes per state-update with no reuse: nctl*64*bV + 2*bV + 1."""
    return nctl * 64 * bv + 2 * bv + 1


# Floating-point operations of the reference's plain backup, counted on its
# source (solver/src/PeregrineSolver.jl): every +, -, *, / and sqrt, with
# comparisons, clamps, floor/round, abs and integer index arithmetic free.
SINCOS_FLOPS = 28          # Float32 sincos: reduction 5, sin kernel 12, cos kernel 11 (double)
SUBSTEP_FLOPS = 140        # diverged 5, body_accel_sat 85, w x v 4, sincos 28, update 18
PROBE_FLOPS = 16           # chord point 7, grid coordinates 6, heading wrap 3
INTERP_FLOPS = 475         # coordinates 12 + wrap 3, fractions 6 + 6, 64 corners x (5 mul + fma 2)


def algorithmic_flops_per_lookahead(nsub: int, checks: int, knee: bool = False) -> int:
    """One `lookahead` of the plain scheme (PeregrineSolver.jl:1100-1153):
    hs = dt/nsub (1) and the command scaled by the traction gain (3, plus
    the gain's sqrt, products and quotients, 10, when a knee is set; the
    tanh is not counted), nsub forward-Euler substeps (:208-221),
    displacement (3), `checks` swept probes (:1117-1135), the successor's
    velocity rotated to the field frame (sincos + 6), the 64-corner
    multilinear read (:399-466) and tau + V (1), min with cap (free)."""
    return 1 + 3 + (10 if knee else 0) + nsub * SUBSTEP_FLOPS + 3 + checks * PROBE_FLOPS + SINCOS_FLOPS + 6 + \
        INTERP_FLOPS + 1


def algorithmic_flops_per_update(nctl: int, nsub: int, checks: int, knee: bool = False) -> int:
    """One state-update of the plain scheme: the cell's coordinates (6 axis
    values, 12), its field velocity rotated into the body frame (sincos + 6),
    and one lookahead per control (`cell_update`, :1205-1344)."""
    return 12 + SINCOS_FLOPS + 6 + nctl * algorithmic_flops_per_lookahead(nsub, checks, knee)


def adjoint_starts(prob: Problem, n: int = 4096) -> np.ndarray:
    """The adjoint objective's start states: x, y uniform over the field,
    heading uniform, velocity and spin zero, redrawn while the start's
    nearest (x, y, heading) cell is blocked; numpy PCG64(4 + 100 * config).
    Rows of (x, y, heading, vx, vy, omega)."""
    r = rng_for(4, prob.cfg)
    g = prob.grid
    occ = np.asarray(prob.occ, np.uint8).reshape(g.n[0], g.n[1], g.n[2])
    out = np.zeros((n, 6), np.float64)
    k = 0
    while k < n:
        x, y = r.uniform(0.0, FIELD), r.uniform(0.0, FIELD)
        h = r.uniform(-math.pi, math.pi)
        ix = int(round((x - g.lo[0]) / g.step[0]))
        iy = int(round((y - g.lo[1]) / g.step[1]))
        ih = int(round((h - g.lo[2]) / g.step[2])) % g.n[2]
        if occ[ix, iy, ih]:
            continue
        out[k, :3] = (x, y, h)
        k += 1
    return out
