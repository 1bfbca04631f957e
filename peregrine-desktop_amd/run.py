"""The solve.jl job interface: config file in, card image out.

Mirrors solver/src/Run.jl and the loaders of solver/src/Pipeline.jl:
`load_model`/`load_field`/`load_targets` (Pipeline.jl:406-500), the grid
from the field, footprint and resolution (`feasible_bounds`,
Pipeline.jl:227-251; `axis_samples`, `symmetric_samples`, `build_grid`,
`grid_inputs`, Run.jl:55-162), every solver knob read once
(`solver_params`, Run.jl:203-294), `check_targets` (Advise.jl:513-570),
the solve driver (`solve_value!`/`solve_escape!`/`run_solve`,
Run.jl:1434-2133) with its PROGRESS stream, MANIFEST.JSON and MODEL.JSON
(`model_json`/`honing_gains`, Run.jl:975-1378), `honing_only`
(Run.jl:1397-1432) and a `plan` report (Run.jl:699-959).

The value solve runs on the device through libperegrine_hip; there is no
CPU backend (`backend: "cpu"` is refused). The tiled (out-of-core) driver
runs the same scheme and the escape pass.
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import sys
import time
import traceback
from dataclasses import dataclass, replace
from datetime import datetime, timezone

import numpy as np

from . import abi, card
from .model import F32, Grid6, Model, Params, control_set, target_cells
from .occupancy import build_occupancy, heading_bin_angles, point_too_close, polys_too_close, wall_box

AXES = ("x", "y", "h", "vx", "vy", "w")
# Run.jl:9-10, 24-27: velocity envelope and resolution defaults.
DEFAULT_VMAX = 170.0
DEFAULT_WMAX = 8.0
DEFAULT_XY_CM = 8.0
DEFAULT_HEADING_DEG = 22.5
DEFAULT_V_CM_S = 34.0
DEFAULT_W_RAD_S = 1.6
# Advise.jl:19, 30, 165: the reference horizon, the measured cost of a
# shorter one (tau seconds, percent of mean value) and the escape pass's
# cell-cost factor used by the time-left estimate.
TAU_REF = 0.5
TAU_COST = ((0.5, 0.0), (0.2, 6.0), (0.1, 21.1), (0.06, 56.0))
ESCAPE_CELL_FACTOR = 1.4


class SolveError(RuntimeError):
    """A configuration or run the solver refuses; the CLI prints it as ERROR."""


def getc(cfg, key, default):
    """cfg[key] unless absent or null (Run.jl:18)."""
    v = cfg.get(key) if isinstance(cfg, dict) else None
    return default if v is None else v


def progress(**kw) -> None:
    """One `PROGRESS {json}` line on stdout, flushed (Run.jl:13-16)."""
    print("PROGRESS " + json.dumps(kw, separators=(",", ":"), default=_jsonable), flush=True)


def _jsonable(v):
    if isinstance(v, np.generic):
        return v.item()
    if isinstance(v, np.ndarray):
        return v.tolist()
    raise TypeError(f"{type(v).__name__} is not JSON serialisable")


def filehash(path: str) -> str:
    """SHA-256 of a file's bytes (Run.jl:2136-2140)."""
    h = hashlib.sha256()
    with open(path, "rb") as fh:
        for blk in iter(lambda: fh.read(1 << 20), b""):
            h.update(blk)
    return h.hexdigest()


def _readjson(path: str) -> dict:
    with open(path, "r", encoding="utf-8") as fh:
        return json.load(fh)


# ---------------------------------------------------------------------------
# Loaders (Pipeline.jl:406-500)
# ---------------------------------------------------------------------------

def load_model(path: str):
    """drivetrain_fit.toml -> (Model, parsed toml). The [mecanum_basis] table
    holds B, A (3x3), q, c (3), optional S, D (3x3) and coulomb_eps (3), and
    a required positive traction_knee (Pipeline.jl:406-440)."""
    import tomli
    with open(path, "rb") as fh:
        cfg = tomli.load(fh)
    if "mecanum_basis" not in cfg:
        raise SolveError(f"{path}: no [mecanum_basis] table; the solve reads its model from a drivetrain fit's mecanum_basis")
    mb = cfg["mecanum_basis"]

    def mat(name):
        if name not in mb:
            return (0.0,) * 9
        return tuple(float(mb[name][r][c]) for r in range(3) for c in range(3))

    if mb.get("traction_knee") is None:
        raise SolveError("the drivetrain fit carries no traction_knee; the gains assume a saturated command, so fit one "
                         "(fit_drivetrain.py --traction-knee auto)")
    knee = F32(mb["traction_knee"])
    if math.isnan(knee) or knee <= 0:
        raise SolveError(f"traction_knee {mb['traction_knee']} is not a positive number; the command saturation needs one")
    eps = tuple(float(v) for v in mb["coulomb_eps"]) if "coulomb_eps" in mb else (5.0, 5.0, 0.15)
    units = cfg.get("units", {}).get("distance", "cm")
    if units != "cm":
        print(f"Warning: regression distance unit is '{units}', solver assumes cm", file=sys.stderr)
    m = Model(mat("B"), mat("A"), tuple(float(v) for v in mb["q"]), mat("S"), mat("D"),
              tuple(float(v) for v in mb["c"]), eps, float(knee))
    return m, cfg


def _poly(pts, what: str) -> np.ndarray:
    if len(pts) < 3:
        raise SolveError(f"{what}: a polygon needs three or more vertices")
    M = np.empty((2, len(pts)), np.float64)
    for i, p in enumerate(pts):
        if len(p) != 2:
            raise SolveError(f"{what}: every vertex has to be an [x, y] pair")
        M[0, i], M[1, i] = float(p[0]), float(p[1])
    return M


def load_field(path: str):
    """field.json -> (bounds (x_min, y_min, x_max, y_max), obstacle polygons,
    robot polygon or None, parsed json) (Pipeline.jl:462-476)."""
    cfg = _readjson(path)
    f = cfg["field"]
    bounds = (float(f["x_min"]), float(f["y_min"]), float(f["x_max"]), float(f["y_max"]))
    polys = [_poly(ob["polygon"], f"obstacle '{ob.get('name', '?')}'") for ob in cfg.get("obstacles", [])]
    robot = None
    if isinstance(cfg.get("robot"), dict) and "polygon" in cfg["robot"]:
        robot = _poly(cfg["robot"]["polygon"], "robot footprint")
    return bounds, polys, robot, cfg


def load_targets(path: str):
    """targets.json -> (names, float32 6-states, parsed json) (Pipeline.jl:478-500)."""
    cfg = _readjson(path)
    names, states = [], []
    for t in cfg["targets"]:
        s = t["state"]
        if len(s) != 6:
            raise SolveError(f"target '{t.get('name', '?')}': a state is six numbers (x, y, heading, vx, vy, omega)")
        names.append(str(t.get("name", f"target{len(names)}")))
        states.append(tuple(F32(v) for v in s))
    if not states:
        raise SolveError(f"{path}: the target list is empty")
    return names, states, cfg


# ---------------------------------------------------------------------------
# Grid (Pipeline.jl:227-251, Run.jl:55-162)
# ---------------------------------------------------------------------------

def feasible_bounds(bounds, robot, clearance: float, nh: int, substeps: int):
    """The x/y box the tracking point can hold at some heading: per heading
    bin the intersection of the wall boxes of its angles, then the union over
    the bins that are possible at all (Pipeline.jl:227-251)."""
    xlo = ylo = math.inf
    xhi = yhi = -math.inf
    nfeas = 0
    for angles in heading_bin_angles(nh, substeps):
        bxlo, bxhi, bylo, byhi = -math.inf, math.inf, -math.inf, math.inf
        for th in angles:
            a, b, c, d = wall_box(bounds, robot, clearance, th)
            bxlo = max(bxlo, a); bxhi = min(bxhi, b)
            bylo = max(bylo, c); byhi = min(byhi, d)
        if not (bxlo <= bxhi and bylo <= byhi):
            continue
        nfeas += 1
        xlo = min(xlo, bxlo); xhi = max(xhi, bxhi)
        ylo = min(ylo, bylo); yhi = max(yhi, byhi)
    if nfeas == 0:
        raise SolveError(
            f"the robot does not fit inside the field at any heading: the field is "
            f"{bounds[2] - bounds[0]:.1f} x {bounds[3] - bounds[1]:.1f} cm, and the footprint plus "
            f"{clearance} cm of clearance needs more than that. Check the units on the robot polygon, "
            f"or lower the clearance")
    return dict(xlo=xlo, xhi=xhi, ylo=ylo, yhi=yhi, headings=nfeas)


def axis_samples(span: float, cell: float):
    """Samples covering `span` at spacing <= `cell`, and the span they cover."""
    if not cell > 0:
        raise SolveError(f"a cell size of {cell} is not positive")
    n = max(2, math.ceil(span / cell - 1.0e-9) + 1)
    return n, (n - 1) * cell


def symmetric_samples(limit: float, cell: float):
    """An odd sample count over [-limit, limit] with a node at zero."""
    if not cell > 0:
        raise SolveError(f"a cell size of {cell} is not positive")
    half = max(1, math.ceil(limit / cell - 1.0e-9))
    return 2 * half + 1, half * cell


def grid_resolution(gcfg):
    r = getc(gcfg, "resolution", {})
    return dict(xy_cm=float(getc(r, "xy_cm", DEFAULT_XY_CM)),
                heading_deg=float(getc(r, "heading_deg", DEFAULT_HEADING_DEG)),
                v_cm_s=float(getc(r, "v_cm_s", DEFAULT_V_CM_S)),
                w_rad_s=float(getc(r, "w_rad_s", DEFAULT_W_RAD_S)))


@dataclass
class GridInputs:
    bounds: tuple
    robot: object
    clearance: float
    wall_clearance: float
    hsub: int


def grid_inputs(cfg, bounds, robot) -> GridInputs:
    """Everything build_grid needs outside the grid block (Run.jl:148-159)."""
    clearance = float(getc(cfg, "clearance_cm", 5.0))
    return GridInputs(bounds, robot, clearance, float(getc(cfg, "wall_clearance_cm", clearance)),
                      int(getc(cfg, "heading_substeps", 3)))


def build_grid(gcfg, gi: GridInputs) -> Grid6:
    """The table's grid (Run.jl:99-138): x/y span the feasible box, padded
    evenly to whole cells; velocity axes symmetric with a node at zero;
    `n` pins the sample counts, `bounds` overrides the inset."""
    vmax = float(getc(gcfg, "vmax", DEFAULT_VMAX))
    wmax = float(getc(gcfg, "wmax", DEFAULT_WMAX))
    res = grid_resolution(gcfg)
    nfix = getc(gcfg, "n", None)
    nh = int(nfix[2]) if nfix is not None else max(4, round(360.0 / res["heading_deg"]))
    ov = getc(gcfg, "bounds", None)
    if ov is not None:
        fb = dict(xlo=float(ov[0]), ylo=float(ov[1]), xhi=float(ov[2]), yhi=float(ov[3]), headings=nh)
    else:
        fb = feasible_bounds(gi.bounds, gi.robot, gi.wall_clearance, nh, gi.hsub)
    if nfix is not None:
        lo = (fb["xlo"], fb["ylo"], -math.pi, -vmax, -vmax, -wmax)
        hi = (fb["xhi"], fb["yhi"], math.pi, vmax, vmax, wmax)
        return Grid6.build(tuple(int(v) for v in nfix), lo, hi)
    nx, sx = axis_samples(fb["xhi"] - fb["xlo"], res["xy_cm"])
    ny, sy = axis_samples(fb["yhi"] - fb["ylo"], res["xy_cm"])
    nv, lv = symmetric_samples(vmax, res["v_cm_s"])
    nw, lw = symmetric_samples(wmax, res["w_rad_s"])
    padx = (sx - (fb["xhi"] - fb["xlo"])) / 2
    pady = (sy - (fb["yhi"] - fb["ylo"])) / 2
    lo = (fb["xlo"] - padx, fb["ylo"] - pady, -math.pi, -lv, -lv, -lw)
    hi = (fb["xhi"] + padx, fb["yhi"] + pady, math.pi, lv, lv, lw)
    return Grid6.build((nx, ny, nh, nv, nv, nw), lo, hi)


def grid_max(g: Grid6, k: int) -> float:
    """The manifest's axis maximum: lo + step*(n-1), or *n on the periodic
    heading, in float32 (Run.jl:2019)."""
    cnt = g.n[k] if k == 2 else g.n[k] - 1
    return float(F32(F32(g.lo[k]) + F32(F32(g.step[k]) * F32(cnt))))


def grid_balance(g: Grid6, m: Model):
    """dv^2/(2a) over the position cell (Run.jl:183-191)."""
    a = max(max(abs(v) for v in m.B[0:3]), max(abs(v) for v in m.B[3:6]))
    if a <= 0:
        return dict(ratio=math.nan, dv_distance_cm=math.nan, accel_cm_s2=0.0)
    dv = float(min(F32(g.step[3]), F32(g.step[4])))
    d = dv * dv / (2 * float(a))
    return dict(ratio=d / float(min(F32(g.step[0]), F32(g.step[1]))), dv_distance_cm=d, accel_cm_s2=float(a))


# ---------------------------------------------------------------------------
# Solver knobs (Run.jl:203-294, 650-680)
# ---------------------------------------------------------------------------

@dataclass
class SolverParams:
    p: Params
    level: int
    nctl: int
    ctl: np.ndarray
    warm: bool
    iters: int
    tol: float
    cap: float
    margin: float
    clearance: float
    ttol: tuple
    tau_auto: bool


def value_cost(tau: float, frm: float | None = None) -> float:
    """Percent added to mean value by a longest step of `tau`, log-linear in
    the measured table; with `frm`, the cost of moving from `frm` down to
    `tau` (Advise.jl:33-50)."""
    if frm is not None:
        return max(0.0, value_cost(tau) - value_cost(frm))
    t = float(tau)
    if t >= TAU_COST[0][0]:
        return 0.0
    for a, b in zip(TAU_COST, TAU_COST[1:]):
        if t >= b[0]:
            f = (math.log(t) - math.log(a[0])) / (math.log(b[0]) - math.log(a[0]))
            return a[1] + f * (b[1] - a[1])
    a, b = TAU_COST[-2], TAU_COST[-1]
    return b[1] + (math.log(t) - math.log(b[0])) * (b[1] - a[1]) / (math.log(b[0]) - math.log(a[0]))


def solver_params(cfg, g: Grid6) -> SolverParams:
    """Read every solver knob once, with the reference's defaults. tau_max
    "auto" (the default) resolves to the reference horizon of 0.5 s, which is
    what it means whenever the grid is held whole."""
    dt = F32(getc(cfg, "dt", 0.05))
    nsub = int(getc(cfg, "substeps", 4))
    tm = getc(cfg, "tau_max", "auto")
    auto = isinstance(tm, str) and tm.lower() == "auto"
    tau_max = float(tm) if isinstance(tm, (int, float)) and not isinstance(tm, bool) else TAU_REF
    hmax = float(getc(cfg, "substep_max", float(F32(dt / F32(max(nsub, 1))))))
    p = Params(dt=float(dt), nsub=nsub, checks=int(getc(cfg, "sweep_checks", 3)),
               adaptive_checks=bool(getc(cfg, "adaptive_checks", True)),
               nearest=bool(getc(cfg, "nearest", False)), cap=float(F32(getc(cfg, "value_cap", 60.0))),
               ntau=int(getc(cfg, "tau_levels", 5)), tau_ratio=float(getc(cfg, "tau_ratio", 2.0)),
               cfl=float(getc(cfg, "cfl", 2.0)), tau_min=float(getc(cfg, "tau_min", 0.004)),
               tau_max=tau_max, hmax=hmax, ncoarse=int(getc(cfg, "control_scan", 8)),
               rounds=int(getc(cfg, "refine_rounds", 2)), delta0=float(getc(cfg, "refine_delta", 0.35)),
               rk2=bool(getc(cfg, "rk2", True)), vclamp=bool(getc(cfg, "velocity_clamp", False)),
               simplex=bool(getc(cfg, "simplex", True)))
    level = int(getc(cfg, "control_level", 4))
    ctl = control_set(level)
    ttol = tuple(F32(v) for v in getc(cfg, "target_tol", [F32(g.step[k]) * F32(0.5) for k in range(6)]))
    return SolverParams(p=p, level=level, nctl=int(ctl.shape[0]), ctl=ctl,
                        warm=bool(getc(cfg, "warm_start", True)), iters=int(getc(cfg, "iterations", 400)),
                        tol=float(getc(cfg, "tolerance", 1e-3)), cap=float(p.cap),
                        margin=float(getc(cfg, "margin_cm", 0.0)), clearance=float(getc(cfg, "clearance_cm", 5.0)),
                        ttol=ttol, tau_auto=auto)


def unsupported_options(sp: SolverParams, ooc: bool) -> list:
    """Scheme settings the chosen driver does not implement, each with the
    setting that avoids it. Both drivers run every setting: the tiled one
    sweeps the pattern search, the horizon ladder and the warm start through
    windows from the scheme's reach, so the list is empty; the plan report
    keeps the field."""
    return []


def quiet_needed(p: Params, nctl: int) -> int:
    """Quiet sweeps in a row before the solve stops: one full turn of the
    rotating control slice (Run.jl:1440-1441)."""
    return 1 if p.ncoarse <= 0 else max(1, -(-(nctl - 1) // p.ncoarse))


# ---------------------------------------------------------------------------
# Targets (Advise.jl:513-582)
# ---------------------------------------------------------------------------

def check_targets(g: Grid6, occ: np.ndarray, names, states, ttol, bounds, polys, robot, clearance: float,
                  wall_clearance: float) -> list:
    """Per target: seed cells, how many are free, the axes it lies off the
    table on, and what blocks it."""
    out = []
    n = g.n
    for i, s in enumerate(states):
        seeds = target_cells(g, s, ttol)
        i1, i2, i3, _, _, _ = g.unflatten(seeds)
        nfree = int(np.count_nonzero(occ[(i1 * n[1] + i2) * n[2] + i3] == 0))
        off = []
        for k in range(6):
            if k == 2:
                continue
            hi = float(g.lo[k]) + float(g.step[k]) * (n[k] - 1)
            if float(s[k]) < float(g.lo[k]) - 1.0e-6 or float(s[k]) > hi + 1.0e-6:
                off.append(AXES[k])
        why = []
        if not off and nfree == 0:
            th = float(s[2])
            bxlo, bxhi, bylo, byhi = wall_box(bounds, robot, wall_clearance, th)
            if float(s[0]) < bxlo or float(s[0]) > bxhi or float(s[1]) < bylo or float(s[1]) > byhi:
                why.append("its footprint reaches past the field edge (wall clearance included)")
            point = robot is None or robot.shape[1] < 3
            body = None
            if not point:
                R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
                body = R @ robot + np.array([[float(s[0])], [float(s[1])]])
            for pi_, P in enumerate(polys, 1):
                hit = point_too_close(float(s[0]), float(s[1]), P, clearance) if point \
                    else polys_too_close(body, P, clearance)
                if hit:
                    why.append(f"it comes closer than {clearance} cm to obstacle {pi_}")
            if not why:
                why.append(f"its own heading is clear but its heading bin is not: a bin of "
                           f"{math.degrees(float(g.step[2])):.1f} deg is blocked when any angle in it is, and "
                           f"finer heading bins would tell the two apart")
        out.append({"index": i, "name": names[i], "seed_cells": int(seeds.size), "free_seed_cells": nfree,
                    "ok": nfree > 0, "off_axes": off, "blocked_by": why})
    return out


def occupancy_summary(g: Grid6, occ: np.ndarray):
    nh = g.n[2]
    o = np.asarray(occ, np.uint8).reshape(-1, nh) != 0
    per = o.mean(axis=0)
    return dict(blocked=int(o.sum()), cells=int(o.size), frac=float(o.mean()),
                worst_heading_frac=float(per.max()), best_heading_frac=float(per.min()))


# ---------------------------------------------------------------------------
# MODEL.JSON and the honing PID (Run.jl:975-1378)
# ---------------------------------------------------------------------------

def inv3(M):
    """3x3 inverse by cofactors; a singular A_u has no controller (Run.jl:975-988)."""
    (a, b, c), (d, e, f), (g, h, i) = M
    det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g)
    if not (math.isfinite(det) and abs(det) > 1e-12):
        raise SolveError(f"the command block of the fit is singular (determinant {det}): the three axes lack independent "
                         f"control authority, and no approach controller exists for such a model")
    return [[(e * i - f * h) / det, (c * h - b * i) / det, (b * f - c * e) / det],
            [(f * g - d * i) / det, (a * i - c * g) / det, (c * d - a * f) / det],
            [(d * h - e * g) / det, (b * g - a * h) / det, (a * e - b * d) / det]]


def _rows(t):
    return [[float(t[3 * r + c]) for c in range(3)] for r in range(3)]


def honing_gains(m: Model, cfg) -> dict:
    """The final-approach PID from the linearised model (Run.jl:1056-1258).

    Near the target the plant is dp/dt = v, dv/dt = Beff*u - Lambda*v with
    Lambda = -(A + S*diag(1/eps)) and Beff = B*tanh(r)/r at the operating
    command r = budget/knee. Kp, Ki, Kd = inv(Beff)*diag(kp, ki, kd), with
    Kd also cancelling Lambda's off-diagonal part, place each decoupled
    channel's poles at -p_i and a double -w: kd_i = p_i + 2w - lambda_i,
    kp_i = w^2 + 2 p_i w, ki_i = p_i w^2, p_i = max(lambda_i - 2w, w) so
    kd_i >= 0. w is the smaller of the bisected saturation bound (Kp times
    the fine band stays inside the budget) and a tenth of the loop rate,
    times bandwidth_scale.
    """
    B, A, S = _rows(m.B), _rows(m.A), _rows(m.S)
    eps = [float(v) for v in m.eps]
    knee = float(m.knee)
    Lam = [[-(A[r][c] + S[r][c] / eps[c]) for c in range(3)] for r in range(3)]
    hcm = float(getc(cfg, "honing_handoff_cm", 15.0))
    hrad = float(getc(cfg, "honing_handoff_rad", 0.25))
    hz = float(getc(cfg, "honing_loop_hz", 50.0))
    frac = float(getc(cfg, "honing_budget_frac", 0.5))
    budget = float(getc(cfg, "honing_budget", frac * knee))
    scale = float(getc(cfg, "honing_bandwidth_scale", 1.0))
    ishare = float(getc(cfg, "honing_integral_share", 0.25))
    if not (budget > 0 and frac > 0 and hz > 0 and scale > 0 and 0 < ishare < 1):
        raise SolveError(f"the honing settings need positive values and an integral share under 1; got "
                         f"honing_budget={budget}, honing_budget_frac={frac}, honing_loop_hz={hz}, "
                         f"honing_bandwidth_scale={scale}, honing_integral_share={ishare}")
    if budget > 1:
        raise SolveError(f"a honing_budget of {budget} exceeds the command set |fwd| + |strafe| + |turn| <= 1, so the "
                         f"mixer would clip what the gains are sized for (honing_budget_frac={frac}, knee {knee})")
    r_op = budget / knee
    tgain = 1.0 if r_op < 1e-6 else math.tanh(r_op) / r_op
    Binv = inv3([[tgain * B[r][c] for c in range(3)] for r in range(3)])
    lam = [Lam[i][i] for i in range(3)]
    Loff = [[0.0 if r == c else Lam[r][c] for c in range(3)] for r in range(3)]
    fcm = float(getc(cfg, "honing_fine_cm", 2.0))
    efine = [fcm, fcm, float(getc(cfg, "honing_fine_rad", 0.03))]

    def gains_at(w):
        p = [max(lam[i] - 2 * w, w) for i in range(3)]
        kp = [w * w + 2 * p[i] * w for i in range(3)]
        ki = [p[i] * w * w for i in range(3)]
        kd = [p[i] + 2 * w - lam[i] for i in range(3)]
        Kp = [[Binv[r][c] * kp[c] for c in range(3)] for r in range(3)]
        Ki = [[Binv[r][c] * ki[c] for c in range(3)] for r in range(3)]
        Kd = [[Binv[r][c] * kd[c] - sum(Binv[r][k] * Loff[k][c] for k in range(3)) for c in range(3)]
              for r in range(3)]
        return Kp, Ki, Kd, p, kd

    def prop_norm(w):
        Kp = gains_at(w)[0]
        return sum(abs(sum(Kp[r][c] * efine[c] for c in range(3))) for r in range(3))

    w_hi = 1e-3
    while prop_norm(w_hi) < budget and w_hi < 1e4:
        w_hi *= 2
    w_lo = w_hi / 2
    for _ in range(80):
        mid = 0.5 * (w_lo + w_hi)
        if prop_norm(mid) < budget:
            w_lo = mid
        else:
            w_hi = mid
    w_sat = w_lo
    w_loop = 2 * math.pi * hz / 10
    w = scale * min(w_sat, w_loop)
    if not (math.isfinite(w) and w > 0):
        raise SolveError(f"the honing bandwidth evaluates to {w}; honing_loop_hz, honing_budget and "
                         f"honing_bandwidth_scale must all be positive")
    Kp, Ki, Kd, p, kd = gains_at(w)
    ilim = []
    for c in range(3):
        col = sum(abs(Ki[r][c]) for r in range(3))
        ilim.append(ishare * budget / (3 * col) if col > 0 else 0.0)
    return {
        "note": "Gains for the PID that takes over from the tables for the final approach, derived from the "
                "same model as the tables. `config` holds what was chosen rather than derived; budget_frac is "
                "the share of the traction knee the approach may spend, which sets the command budget and the "
                "bandwidth. Additive to schema 2.",
        "frame": "robot body",
        "law": "u = Kp*e + Ki*clamp(integral(e), -integral_limit, integral_limit) + Kd*d(e)/dt + u_ff",
        "error": {"vector": ["e_fwd", "e_strafe", "e_turn"], "units": ["cm", "cm", "rad"],
                  "note": "e = target - current, rotated into the ROBOT BODY frame (rotate by -h; the tables "
                          "are field frame)."},
        "feedforward": {"formula": "u_ff = B_eff_inv*(a_ref + Lambda*v_ref)",
                        "why": "Model inversion: the command the fit says produces the reference motion. "
                               "Optional; Lambda already holds the linearised Coulomb term."},
        "Kp": Kp, "Ki": Ki, "Kd": Kd, "B_eff_inv": Binv, "Lambda": Lam,
        "traction_gain": {"value": tgain,
                          "why": f"B_eff = A_u * tanh(r)/r at r = budget/knee = {round(r_op, 4)}; B_eff_inv "
                                 f"already contains it. Still saturate the final command."},
        "integral_limit": {"value": ilim, "units": ["cm*s", "cm*s", "rad*s"],
                           "why": f"Clamp each integrator to this before multiplying by Ki, so the integral "
                                  f"term claims at most {round(ishare * 100)} percent of the command budget; "
                                  f"stop integrating while the command is saturated."},
        "omega": w,
        "poles": {"double": w, "third": p, "kd_diag": kd,
                  "why": "Per axis the closed loop is (s+third)(s+omega)^2; kd is never negative."},
        "omega_limits": {"saturation": w_sat, "loop_rate": w_loop,
                         "bound_by": "saturation" if w_sat <= w_loop else "loop_rate",
                         "why": "The smaller sets the bandwidth: Kp*e stays inside the budget over the fine "
                                "band, and the poles stay within a tenth of the sampling rate."},
        "handoff": {"cm": hcm, "rad": hrad,
                    "note": "Where the tables hand over; engaging further out saturates the command longer."},
        "fine_band": {"cm": efine[0], "rad": efine[2],
                      "why": "Inside this band the command stays within the budget."},
        "config": {"loop_hz": hz, "budget": budget, "budget_frac": budget / knee, "bandwidth_scale": scale,
                   "integral_share": ishare, "fine_cm": efine[0], "fine_rad": efine[2]},
        "derivation": "TABLE_FORMAT.md section 8.8",
    }


def model_json(m: Model, reg_path: str, zeroed_c: bool, cfg=None) -> dict:
    """MODEL.JSON, schema 2 (Run.jl:1278-1378): the body-frame model as
    a = A_s*s + A_u*u + A_ss*(s.*s) + A_uu*(u.*u) + A_sgn*csign(s)
    + A_absv*(abs(s).*s) + k over the 6-state, plus the honing block."""
    cfg = cfg or {}
    A, B, S, D = _rows(m.A), _rows(m.B), _rows(m.S), _rows(m.D)
    A_s = [[0.0, 0.0, 0.0] + A[r] for r in range(3)]
    A_ss = [[0.0, 0.0, 0.0, 0.0, 0.0, float(m.q[r])] for r in range(3)]
    A_u = B
    A_uu = [[0.0, 0.0, 0.0] for _ in range(3)]
    A_sgn = [[0.0, 0.0, 0.0] + S[r] for r in range(3)]
    A_abs = [[0.0, 0.0, 0.0] + D[r] for r in range(3)]
    k = [float(v) for v in m.c]
    blocks = (("A_s", A_s), ("A_u", A_u), ("A_ss", A_ss), ("A_uu", A_uu), ("A_sgn", A_sgn), ("A_absv", A_abs))
    nz = [nm for nm, blk in blocks if any(v != 0 for r in blk for v in r)]
    if any(v != 0 for v in k):
        nz.append("k")
    return {
        "schema_version": 2,
        "generator": "peregrine-desktop_amd",
        "regression_sha256": filehash(reg_path),
        "equation": "a = A_s*s + A_u*u + A_ss*(s.*s) + A_uu*(u.*u) + A_sgn*csign(s) + A_absv*(abs(s).*s) + k",
        "csign": {"formula": "csign(s)_i = clamp(s_i / coulomb_eps_i, -1, 1)",
                  "coulomb_eps": [0.0, 0.0, 0.0] + [float(v) for v in m.eps],
                  "why": "Coulomb friction smoothed instead of a hard sign(). Use exactly this form."},
        "output": {"vector": ["a_x", "a_y", "alpha"], "units": ["cm/s^2", "cm/s^2", "rad/s^2"],
                   "frame": "robot body"},
        "state": {"vector": list(AXES), "units": ["cm", "cm", "rad", "cm/s", "cm/s", "rad/s"],
                  "frame_note": "IMPORTANT: vx and vy must be rotated into the ROBOT BODY frame before use "
                                "here. The value tables index field-frame velocity, so rotate by -h between "
                                "the two. x, y and h have zero coefficients throughout."},
        "control": {"vector": ["fwd", "strafe", "turn"],
                    "note": "mecanum projection of the wheel powers; the admissible set is "
                            "|fwd| + |strafe| + |turn| <= 1",
                    "saturation": {"knee": float(m.knee),
                                   "formula": "m = norm(u_raw); u = u_raw * tanh(m/knee) / (m/knee)",
                                   "why": "Apply before A_u: the gains were fitted against the saturated "
                                          "command. The knee is always present and positive."}},
        "A_s": A_s, "A_u": A_u, "A_ss": A_ss, "A_uu": A_uu, "A_sgn": A_sgn, "A_absv": A_abs, "k": k,
        "constant_zeroed": bool(zeroed_c),
        "nonzero_blocks": nz,
        "honing": honing_gains(m, cfg),
    }


def _model_for(cfg):
    reg_path = str(cfg["regression"])
    m, regcfg = load_model(reg_path)
    zeroed = bool(getc(cfg, "zero_c", True))
    if zeroed:
        m = replace(m, c=(0.0, 0.0, 0.0))
    return m, reg_path, zeroed


def honing_only(cfg, write_dir=None) -> dict:
    """The honing block from the regression alone; with `write_dir`, rebuild
    that run's MODEL.JSON in place, refused when it was written from a
    different regression (Run.jl:1397-1432)."""
    m, reg_path, zeroed = _model_for(cfg)
    mj = model_json(m, reg_path, zeroed, cfg)
    if write_dir is not None:
        path = os.path.join(str(write_dir), "MODEL.JSON")
        if not os.path.isfile(path):
            raise SolveError(f"{write_dir} holds no MODEL.JSON to update; run the solve that writes it first")
        old = _readjson(path)
        if str(getc(old, "regression_sha256", "")) != mj["regression_sha256"]:
            raise SolveError(f"{path} belongs to tables solved from a different regression than {reg_path}; gains "
                             f"from this fit would not match them")
        with open(path, "w") as fh:
            json.dump(mj, fh, indent=2)
    return mj["honing"]


# ---------------------------------------------------------------------------
# Plan (Run.jl:699-959)
# ---------------------------------------------------------------------------

def _setup(cfg):
    bounds, polys, robot, _ = load_field(str(cfg["field"]))
    gcfg = getc(cfg, "grid", {})
    gi = grid_inputs(cfg, bounds, robot)
    g = build_grid(gcfg, gi)
    return bounds, polys, robot, gcfg, gi, g


def _device_budget(headroom: float):
    """Free device bytes times headroom, or None without a device."""
    from . import solver
    try:
        return solver.gpu_budget(headroom)
    except (abi.PeregrineError, OSError):
        return None


def _settle(cfg, g: Grid6, m: Model, sp: SolverParams) -> dict:
    """settle + decompose (Run.jl:316-354, 650-682) with this build's device
    budget, and the solver parameters with the resolved tau_max."""
    from . import tiling
    st = tiling.settle(cfg, g, m, sp.p, sp.ctl, sp.nctl, sp.warm, sp.tau_auto, _device_budget)
    st["sp"] = replace(sp, p=st["p"])
    return st


def ooc_report(cfg, g: Grid6, dec: dict, iters: int, ntargets: int) -> dict:
    """PLAN's out_of_core block (Run.jl:721-781): the mode, and when tiled
    either the tile plan (halo, tile, loaded extent, tiles per round,
    resident bytes, amplification, reach, store bytes, bytes moved per round)
    or, when no tiling fits, the advice that would make one fit."""
    from . import tiling
    ooc = {"mode": dec["mode"], "budget_bytes": dec["budget"], "whole_bytes": dec["whole_bytes"]}
    if dec["mode"] != "ooc":
        return ooc
    cell_cm = min(float(g.step[0]), float(g.step[1]))
    tp = dec["tp"]
    if tp is None:
        a = dec["advice"]
        ooc.update(fits=False, reach_cm=dec["dxy_cm"], reach_cells=dec["dxy_cm"] / cell_cm,
                   halo_cells=[dec["hx"], dec["hy"]], column_bytes=a["column_bytes"],
                   max_halo_cells=a["max_halo"], min_window_cells=1 + 2 * max(dec["hx"], dec["hy"]),
                   warm_start_would_help=a["warm_helps"])
        rounds = max(1, iters)
        for pre, o in (("suggest", a["tau_max"]), ("comfortable", a["tau_comfortable"])):
            if o is None:
                continue
            ooc.update({pre + "_tau_max": o["tau"], pre + "_halo_cells": o["halo"], pre + "_reach_cm": o["reach_cm"],
                        pre + "_tile_cells": o["tile"], pre + "_amplification": o["amplification"],
                        pre + "_value_cost": o["cost"], pre + "_total_s": o["round_s"] * rounds * ntargets})
        return ooc
    wt = bool(getc(cfg, "warm_start_tiled", True))
    sb = tiling.cell_bytes(wt)
    store = g.ncells * sb
    per_round = int(round(dec["col"] * sb * tp["nxl"] * tp["nyl"] * -(-g.n[0] // tp["wx"]) *
                          -(-g.n[1] // tp["wy"]))) + store
    ooc.update(fits=True, halo_cells=[tp["hx"], tp["hy"]], tile_cells=[tp["wx"], tp["wy"]],
               loaded_cells=[tp["nxl"], tp["nyl"]], tiles_per_round=tp["ntiles"], resident_bytes=tp["bytes"],
               amplification=tp["amplification"], reach_cm=tp["dxy_cm"], reach_rad=tp["dh_rad"],
               reach_cells=tp["dxy_cm"] / cell_cm, store_bytes=store, io_bytes_per_round=per_round, tile_sweeps=1,
               store_in="pinned host memory")
    return ooc


def _device_fit(g: Grid6):
    """(available, fits whole, need GiB, free GiB) from the library."""
    from . import solver
    try:
        fits, need, free = solver.gpu_fits(g)
        return True, fits, need, free
    except (abi.PeregrineError, OSError):
        return False, False, g.ncells * 4 / 2 ** 30 * 2, 0.0


# Seconds per lookahead on one MI355X, measured: the global-gather kernel
# running the default scheme (1.521 s for 564,854,784 states of at most 27
# lookaheads each) and the LDS-staged kernel running the plain scheme (1.532 s
# for the same states x 125 lookaheads). The escape pass costs its share of a
# value sweep (0.059 s against 0.107 s on a 35.8M-state default job).
LOOKAHEAD_S = {"global": 1.521 / (564854784 * 27), "lds": 1.532 / (564854784 * 125)}
ESCAPE_SHARE = 0.059 / 0.107


def lookaheads_per_state(p: Params, nctl: int, warm: bool) -> int:
    """The most lookaheads one backup evaluates: coast, the warm start, the
    lattice or its slice, the search moves, the ladder rungs and dt rung."""
    nnc = max(nctl - 1, 1)
    span = nnc if p.ncoarse <= 0 else min(p.ncoarse, nnc)
    n = 1 + int(warm) + span + 6 * max(p.rounds, 0)
    if p.ntau > 1:
        n += p.ntau - 1 + (1 if p.cfl > 0 else 0)
    return n


def runtime_estimate(g: Grid6, sp: SolverParams, esc_iters: int, ntargets: int) -> dict:
    """An upper bound on the run's time from the sweep budget and the
    measured per-lookahead cost of the kernel the scheme runs on (the
    iteration budget; the tolerance usually stops sooner)."""
    p = sp.p
    kernel = "global" if (p.rounds > 0 or p.ntau > 1 or sp.warm) else "lds"
    unit_s = g.ncells * lookaheads_per_state(p, sp.nctl, sp.warm) * LOOKAHEAD_S[kernel]
    units = sp.iters + esc_iters
    per_target = sp.iters * unit_s + esc_iters * unit_s * ESCAPE_SHARE
    return {"unit": "sweep", "unit_s": unit_s, "units": units, "escape_sweeps": esc_iters,
            "total_s": per_target * ntargets, "kernel": kernel,
            "lookaheads_per_state": lookaheads_per_state(p, sp.nctl, sp.warm)}


def suggest_heading_deg(reach_cm: float, xy_cm: float) -> float:
    """Heading bins such that the footprint's outer point moves about one
    position cell per bin: 2 pi reach / xy bins, clamped to [8, 180]
    (Advise.jl:361-365)."""
    r = max(float(reach_cm), 1.0e-6)
    nb = min(max(math.ceil(2 * math.pi * r / float(xy_cm)), 8), 180)
    return 360.0 / nb


def suggest_resolution(span_x: float, span_y: float, reach_cm: float, vmax: float, wmax: float, elem_bytes: int,
                       ntargets: int, budget_bytes: float, finest_cm: float = 1.0, vram_bytes: float = math.inf,
                       bytes_per_cell: int = 16):
    """The finest position cell (from finest_cm, 0.1 cm steps below 10 cm,
    0.5 cm above, up to 60 cm) whose card bytes fit budget_bytes and whose
    device bytes fit vram_bytes; heading from the footprint, velocity and
    spin a tenth of their envelope (21 samples each). bound_by says which
    budget stopped it (Advise.jl:400-434)."""
    v_cm_s, w_rad_s = vmax / 10, wmax / 10
    nv = nw = 21
    xy = float(finest_cm)
    while xy <= 60.0:
        hdeg = suggest_heading_deg(reach_cm, xy)
        nh = max(4, round(360.0 / hdeg))
        n1 = max(2, math.ceil(span_x / xy - 1.0e-9) + 1)
        n2 = max(2, math.ceil(span_y / xy - 1.0e-9) + 1)
        cells = n1 * n2 * nh * nv * nv * nw
        nbytes = float(cells) * elem_bytes * max(ntargets, 1)
        vram = float(cells) * bytes_per_cell
        if nbytes <= budget_bytes and vram <= vram_bytes:
            return {"xy_cm": xy, "heading_deg": hdeg, "v_cm_s": v_cm_s, "w_rad_s": w_rad_s,
                    "n": [n1, n2, nh, nv, nv, nw], "cells": cells, "bytes": nbytes, "vram_bytes": vram,
                    "bound_by": "vram" if vram > 0.5 * vram_bytes else "card"}
        xy += 0.1 if xy < 10 else 0.5
    return None


def shape_for(span_x: float, span_y: float, vmax: float, wmax: float, xy: float, hdeg: float, v: float, w: float):
    """The sample counts build_grid gives a resolution (Advise.jl:436-447)."""
    n1 = max(2, math.ceil(span_x / xy - 1.0e-9) + 1)
    n2 = max(2, math.ceil(span_y / xy - 1.0e-9) + 1)
    nh = max(4, round(360.0 / hdeg))
    nv = 2 * max(1, math.ceil(vmax / v - 1.0e-9)) + 1
    nw = 2 * max(1, math.ceil(wmax / w - 1.0e-9)) + 1
    return [n1, n2, nh, nv, nv, nw], n1 * n2 * nh * nv * nv * nw


def coarsen_to_fit(span_x: float, span_y: float, vmax: float, wmax: float, res: dict, budget_bytes: float,
                   bytes_per_cell: int = 16):
    """The smallest common scale (1 to 16, steps of 0.05) of all four cell
    sizes at which the grid holds whole on the device (Advise.jl:469-511)."""
    k = 1.0
    while k <= 16.0:
        n, cells = shape_for(span_x, span_y, vmax, wmax, res["xy_cm"] * k, res["heading_deg"] * k,
                             res["v_cm_s"] * k, res["w_rad_s"] * k)
        if float(cells) * bytes_per_cell <= budget_bytes:
            return {"scale": k, "xy_cm": res["xy_cm"] * k, "heading_deg": res["heading_deg"] * k,
                    "v_cm_s": res["v_cm_s"] * k, "w_rad_s": res["w_rad_s"] * k, "n": n, "cells": cells,
                    "vram_bytes": float(cells) * bytes_per_cell}
        k += 0.05
    return None


def device_bytes_per_cell(warm: bool) -> int:
    """Device bytes per state of a whole-grid FP32 solve: the Jacobi pair of
    tables and, with the warm start, its 3-byte store."""
    return 8 + (3 if warm else 0)


def plan(cfg) -> dict:
    """Size and feasibility report: grid, bounds, resolution, occupancy,
    target checks, bytes on the card, whether the table fits the device, the
    horizon in force, a run-time bound from measured kernel costs, and the
    resolution advice (the finest resolution within the card and device
    budgets; the common coarsening that holds the grid whole). The reference's
    horizon/tiling advice for its own tiled driver is not part of it."""
    bounds, polys, robot, gcfg, gi, g = _setup(cfg)
    m, _, _ = _model_for(cfg)
    names, states, _ = load_targets(str(cfg["targets"]))
    sp = solver_params(cfg, g)
    dtype = str(getc(cfg, "dtype", "u16"))
    if dtype not in card.DTYPES:
        raise SolveError(f"dtype '{dtype}' is not one of {', '.join(card.DTYPES)}")
    chunk_elements = int(getc(cfg, "chunk_elements", 1 << 23))
    st = _settle(cfg, g, m, sp)
    sp, dec, rtau = st["sp"], st["dec"], st["rec"]
    occ = build_occupancy(g, polys, sp.margin, robot, gi.hsub, sp.clearance, bounds=bounds,
                          wall_clearance_cm=gi.wall_clearance)
    osum = occupancy_summary(g, occ)
    tchk = check_targets(g, occ, names, states, sp.ttol, bounds, polys, robot, sp.clearance, gi.wall_clearance)
    cells = g.ncells
    eb = card.DTYPES[dtype][0]
    avail, fits, need, free = _device_fit(g)
    res = grid_resolution(gcfg)
    # resolution advice (Run.jl:800-833): a size budget on the card, the
    # footprint's radius for the heading rule, and the device's free bytes
    rec_budget = float(getc(cfg, "size_budget_bytes", 8 * 2 ** 30))
    reach_cm = 0.0 if robot is None or robot.shape[1] < 3 else float(np.max(np.hypot(robot[0], robot[1])))
    vmax = float(getc(gcfg, "vmax", DEFAULT_VMAX))
    wmax = float(getc(gcfg, "wmax", DEFAULT_WMAX))
    bpc = device_bytes_per_cell(sp.warm)
    dev_budget = float(dec["budget"]) if dec["budget"] > 0 else 0.0
    span_x = float(g.step[0]) * (g.n[0] - 1)
    span_y = float(g.step[1]) * (g.n[1] - 1)
    sugg = suggest_resolution(span_x, span_y, reach_cm, vmax, wmax, card.DTYPES[dtype][0], len(names), rec_budget,
                              vram_bytes=dev_budget if dev_budget > 0 else math.inf, bytes_per_cell=bpc)
    incore = None
    if dec["mode"] == "ooc" and dev_budget > 0:
        incore = coarsen_to_fit(span_x, span_y, vmax, wmax,
                                {"xy_cm": float(min(g.step[0], g.step[1])),
                                 "heading_deg": math.degrees(float(g.step[2])), "v_cm_s": float(g.step[3]),
                                 "w_rad_s": float(g.step[5])}, dev_budget, bytes_per_cell=bpc)
    bal = grid_balance(g, m)
    full = (bounds[2] - bounds[0]) * (bounds[3] - bounds[1])
    span = (grid_max(g, 0) - float(g.lo[0])) * (grid_max(g, 1) - float(g.lo[1]))
    return {
        "cells": cells, "axes": list(AXES), "n": list(g.n),
        "bounds": {
            "field": list(bounds),
            "table_min": [float(v) for v in g.lo],
            "table_max": [grid_max(g, k) for k in range(6)],
            "inset_cm": [float(g.lo[0]) - bounds[0], float(g.lo[1]) - bounds[1],
                         bounds[2] - grid_max(g, 0), bounds[3] - grid_max(g, 1)],
            "cells_saved_frac": 1.0 - span / full if full > 0 else 0.0,
            "wall_clearance_cm": gi.wall_clearance, "clearance_cm": gi.clearance,
        },
        "resolution": {
            "requested": res,
            "actual": {"xy_cm": float(min(g.step[0], g.step[1])), "heading_deg": math.degrees(float(g.step[2])),
                       "v_cm_s": float(g.step[3]), "w_rad_s": float(g.step[5])},
            "pinned_by_n": getc(gcfg, "n", None) is not None,
        },
        "occupancy": {"blocked_frac": osum["frac"], "blocked_cells": osum["blocked"], "xyh_cells": osum["cells"],
                      "worst_heading_blocked_frac": osum["worst_heading_frac"],
                      "best_heading_blocked_frac": osum["best_heading_frac"]},
        "targets": tchk, "targets_ok": all(t["ok"] for t in tchk),
        "dtype": dtype, "elem_bytes": eb, "bytes_per_target": cells * eb, "n_targets": len(names),
        "bytes_total": cells * eb * len(names), "chunks_per_target": -(-cells // chunk_elements),
        "gpu_available": avail, "gpu_fits": bool(fits), "gpu_need_gb": need, "gpu_free_gb": free,
        "out_of_core": ooc_report(cfg, g, dec, sp.iters, len(names)),
        "grid_balance": bal["ratio"], "dv_distance_cm": bal["dv_distance_cm"],
        "cell_cm": float(min(g.step[0], g.step[1])), "cell_cm_s": float(min(g.step[3], g.step[4])),
        "controls": sp.nctl, "unsupported_options": unsupported_options(sp, dec["mode"] == "ooc"),
        "recommend": {
            "size_budget_bytes": rec_budget, "resolution": sugg, "in_core": incore,
            "tau_applied": float(sp.p.tau_max), "tau_applied_cost": value_cost(sp.p.tau_max, TAU_REF),
            "tau_auto": sp.tau_auto, "tau_reference": TAU_REF,
            "tau_pinned_waste": (not sp.tau_auto) and dec["mode"] != "ooc" and sp.p.tau_max < TAU_REF,
            "tau_max": rtau["tau_max"] if rtau else None, "tau_value_cost": rtau["cost"] if rtau else None,
            "tau_current": rtau["current"] if rtau else float(sp.p.tau_max),
            "tau_current_cost": rtau["current_cost"] if rtau else None,
            "tau_io_minor": rtau["io_minor"] if rtau else None,
            "tau_reason": rtau["reason"] if rtau else None,
            "tau_options": [{k: o[k] for k in ("tau", "cost", "round_s", "compute_s", "io_s", "amplification",
                                               "reach_cm")} for o in (rtau["options"] if rtau else [])],
        },
        "runtime": runtime_estimate(g, sp, int(getc(cfg, "escape_iterations", 60))
                                    if bool(getc(cfg, "escape", True)) else 0, len(names)),
    }


# ---------------------------------------------------------------------------
# Solve (Run.jl:1434-2133)
# ---------------------------------------------------------------------------

def solve_value(s, seeds, p: Params, nctl: int, iters: int, tol: float, on_progress=None):
    """solve_value!: cap everywhere, seeds zero, sweeps alternating
    direction with the sweep index as the slice phase, stopped after
    quiet_needed sweeps in a row whose summed decrease is <= tol
    (Run.jl:1434-1474). Returns (iterations done, last summed decrease)."""
    s.fill_seed(seeds)
    need = quiet_needed(p, nctl)
    last, done, quiet = math.inf, 0, 0
    for it in range(1, iters + 1):
        r = s.sweep(it - 1, it % 2 == 1)
        done, last = it, r.sum_decrease
        quiet = quiet + 1 if last <= tol else 0
        if on_progress is not None and (it % 5 == 0 or it == 1 or quiet >= need):
            on_progress(it, last)
        if quiet >= need:
            break
    return done, last


def solve_escape(s, p: Params, nctl: int, iters: int, tol: float, on_progress=None):
    """solve_escape!: escape sweeps from the converged table, same quiet rule
    (Run.jl:1503-1530)."""
    need = quiet_needed(p, nctl)
    last, done, quiet = math.inf, 0, 0
    for it in range(1, iters + 1):
        r = s.escape_sweep(it - 1, it % 2 == 1)
        done, last = it, r.sum_decrease
        quiet = quiet + 1 if last <= tol else 0
        if on_progress is not None and (it % 5 == 0 or it == 1 or quiet >= need):
            on_progress(it, last)
        if quiet >= need:
            break
    return done, last


def _steady(el, first, done):
    return max(el - first, 0.0) / (done - 1) if done > 1 else el


def run_solve(cfg) -> dict:
    """Solve every target and write the card image under out_dir: TABLES/,
    MANIFEST.JSON, MODEL.JSON (Run.jl:1539-2133)."""
    from . import solver
    reg_path, field_path, targ_path = str(cfg["regression"]), str(cfg["field"]), str(cfg["targets"])
    out_dir = str(cfg["out_dir"])
    m, _, zeroed = _model_for(cfg)
    bounds, polys, robot, gcfg, gi, g = _setup(cfg)
    names, states, _ = load_targets(targ_path)
    cells = g.ncells
    dtype = str(getc(cfg, "dtype", "u16"))
    if dtype not in card.DTYPES:
        raise SolveError(f"dtype '{dtype}' is not one of {', '.join(card.DTYPES)}")
    scale = float(getc(cfg, "scale", card.DTYPES[dtype][2]))
    chunk_elements = int(getc(cfg, "chunk_elements", 1 << 23))
    if chunk_elements <= 0 or chunk_elements & (chunk_elements - 1):
        raise SolveError(f"chunk_elements {chunk_elements} is not a power of two")
    sp = solver_params(cfg, g)
    cap = sp.cap
    backend = str(getc(cfg, "backend", "auto"))
    if backend not in ("auto", "cuda", "hip"):
        raise SolveError(f"backend '{backend}' is not built: this solver runs the sweep on the GPU only")
    do_escape = bool(getc(cfg, "escape", True))
    esc_iters = int(getc(cfg, "escape_iterations", 60))
    esc_tol = float(getc(cfg, "escape_tolerance", sp.tol))
    escape_scale = card.escape_scale_for(dtype, cap) if do_escape else 0.0
    ebase = card.DTYPES[dtype][3]
    if do_escape and ebase > 0 and cap / scale >= ebase:
        raise SolveError(f"with scale {scale}, value_cap {cap} s encodes as {round(cap / scale)}, inside the escape "
                         f"codes that start at {ebase}: keep value_cap under {round(ebase * scale, 2)} s, coarsen "
                         f"the scale, or turn the escape pass off")
    progress(phase="setup", cells=cells, n=list(g.n), controls=sp.nctl, dtype=dtype, escape=do_escape,
             escape_scale=escape_scale, escape_base=ebase, escape_codes=card.escape_codes(dtype),
             bytes_per_target=cells * card.DTYPES[dtype][0], iterations=sp.iters,
             escape_iterations=esc_iters if do_escape else 0, n_targets=len(names))
    occ = build_occupancy(g, polys, sp.margin, robot, gi.hsub, sp.clearance, bounds=bounds,
                          wall_clearance_cm=gi.wall_clearance)
    blocked = int(np.count_nonzero(occ))
    progress(phase="occupancy", blocked_cells=blocked, blocked_frac=blocked / occ.size,
             robot_vertices=0 if robot is None else int(robot.shape[1]), heading_substeps=gi.hsub,
             clearance_cm=sp.clearance, wall_clearance_cm=gi.wall_clearance)
    tchk = check_targets(g, occ, names, states, sp.ttol, bounds, polys, robot, sp.clearance, gi.wall_clearance)
    badt = [t for t in tchk if not t["ok"]]
    if badt:
        lines = []
        for t in badt:
            if t["off_axes"]:
                span = ", ".join(f"{AXES[k]} {float(g.lo[k]):.1f}..{float(g.lo[k]) + float(g.step[k]) * (g.n[k] - 1):.1f}"
                                 for k in (0, 1, 3, 4, 5))
                lines.append(f"  '{t['name']}' lies off the table on {', '.join(t['off_axes'])}; the table "
                             f"covers {span}")
            else:
                lines.append(f"  '{t['name']}' " + "; ".join(t["blocked_by"]))
        raise SolveError("no seed cell of these targets is a state the robot can be in, so every table entry "
                         "would come out unreachable:\n" + "\n".join(lines))
    st = _settle(cfg, g, m, sp)
    sp, dec = st["sp"], st["dec"]
    p = sp.p
    if dec["mode"] == "cpu":
        raise SolveError("no HIP device is visible and no vram_budget_bytes is set: this solver runs the sweep "
                         "on the GPU only")
    ooc = dec["mode"] == "ooc"
    bad = unsupported_options(sp, ooc)
    if bad:
        raise SolveError("the grid does not fit the device whole, and the tiled driver does not run "
                         + ", ".join(bad))
    progress(phase="backend", backend="hip_tiled" if ooc else "hip")
    tables_dir = os.path.join(out_dir, "TABLES")
    os.makedirs(tables_dir, exist_ok=True)
    blocked_frac = blocked / max(occ.size, 1)
    esc_ratio = ESCAPE_CELL_FACTOR * blocked_frac / max(1.0 - blocked_frac, 1.0e-6)
    esc_units = esc_iters if do_escape else 0
    ntargets = len(states)

    def run_eta(ti, solve_done, solve_unit, esc_done, esc_unit):
        return (max(0.0, (sp.iters - solve_done) * solve_unit) + max(0.0, (esc_units - esc_done) * esc_unit)
                + (ntargets - ti) * (sp.iters * solve_unit + esc_units * esc_unit))

    entries = []
    # tile_cells (this build's key): the tiled driver's interior (x, y,
    # heading cells); 0s let the budget choose the largest that fits
    tile = [int(v) for v in getc(cfg, "tile_cells", [0, 0, 0])]
    if len(tile) != 3 or min(tile) < 0:
        raise SolveError(f"tile_cells must be three non-negative cell counts (x, y, heading), got {tile}")
    kw = dict(ooc_budget=int(dec["budget"]), ooc_tile=tuple(tile)) if ooc else {}
    tiling = None
    with solver.HipSolver(g, m, p, sp.ctl, occ, **kw) as s:
        if sp.warm:
            s.set_warm_start(True)
        if ooc:
            tile, ntiles, dev_bytes, amp = s.ooc_info()
            tiling = {"tile_cells": list(tile), "tiles_per_round": ntiles, "tile_sweeps": 1,
                      "device_bytes": dev_bytes, "budget_bytes": int(dec["budget"]), "amplification": amp,
                      "windows": "each tile's exact read range", "warm_start_tiled": sp.warm,
                      "store": "pinned host memory"}
            progress(phase="tiling", **tiling)
        for ti0, st in enumerate(states):
            ti = ti0 + 1
            seeds = target_cells(g, st, sp.ttol)
            t0 = time.time()
            first = [0.0]

            def on_solve(it, d):
                el = time.time() - t0
                if it == 1:
                    first[0] = el
                per = _steady(el, first[0], it)
                progress(phase="solve", target=ti0, target_name=names[ti0], iter=it, iters=sp.iters, delta=d,
                         elapsed_s=el, sweep_s=per,
                         eta_s=run_eta(ti, it, per, 0, esc_ratio * per) if it > 1 else None)

            done, last = solve_value(s, seeds, p, sp.nctl, sp.iters, sp.tol, on_solve)
            solve_unit = _steady(time.time() - t0, first[0], max(done, 1))
            esc_done, esc_delta = 0, 0.0
            if do_escape:
                te = time.time()
                efirst = [0.0]

                def on_escape(it, d):
                    el = time.time() - te
                    if it == 1:
                        efirst[0] = el
                    per = _steady(el, efirst[0], it)
                    progress(phase="escape", target=ti0, target_name=names[ti0], iter=it, iters=esc_iters,
                             delta=d, elapsed_s=el, sweep_s=per,
                             eta_s=run_eta(ti, sp.iters, solve_unit, it, per) if it > 1 else None)

                esc_done, esc_delta = solve_escape(s, p, sp.nctl, esc_iters, esc_tol, on_escape)
            progress(phase="encode", target=ti0, target_name=names[ti0], iters_done=done, delta=last)
            V = s.download()
            info = card.write_table(tables_dir, ti0, V, dtype, scale, chunk_elements, cap, escape_scale)
            del V
            reached = info.reached / info.cells
            escf = info.escaped / info.cells
            dead = info.cells - info.reached
            eou = info.escaped / dead if dead > 0 else 0.0
            entries.append(card_entry(ti0, names[ti0], st, info, done, last, esc_done, esc_delta))
            progress(phase="target_done", target=ti0, target_name=names[ti0], chunks=info.n_chunks,
                     bytes=info.bytes, reached_frac=reached, escape_frac=escf, escape_of_unreached_frac=eou,
                     elapsed_s=time.time() - t0)
    man, mj = write_card_files(out_dir, cfg, g, gi, bounds, sp, dtype, scale, chunk_elements, do_escape, esc_iters,
                               esc_tol, escape_scale, robot, ooc, entries, reg_path, field_path, m, zeroed, tiling)
    progress(phase="model", file="MODEL.JSON", nonzero_blocks=mj["nonzero_blocks"],
             honing_omega=mj["honing"]["omega"], honing_budget_frac=mj["honing"]["config"]["budget_frac"],
             honing_bound_by=mj["honing"]["omega_limits"]["bound_by"])
    progress(phase="done", out_dir=out_dir, targets=len(entries))
    return man


def card_entry(ti0: int, name: str, st, info, done: int, last: float, esc_done: int, esc_delta: float) -> dict:
    """One target's MANIFEST.JSON entry from its written table (Run.jl:2040-2060)."""
    reached = info.reached / info.cells
    dead = info.cells - info.reached
    return {"index": ti0, "name": name, "state": [float(v) for v in st],
            "file_pattern": f"TABLES/T{ti0:02d}C%04d.BIN", "n_chunks": info.n_chunks, "bytes": info.bytes,
            "sha256": info.sha256, "iterations": done, "final_delta": last, "reached_frac": reached,
            "escape_frac": info.escaped / info.cells,
            "escape_of_unreached_frac": info.escaped / dead if dead > 0 else 0.0, "escape_iterations": esc_done,
            "escape_final_delta": esc_delta}


def write_card_files(out_dir, cfg, g, gi, bounds, sp, dtype, scale, chunk_elements, do_escape, esc_iters, esc_tol,
                     escape_scale, robot, ooc, entries, reg_path, field_path, m, zeroed, tiling=None):
    """MANIFEST.JSON and MODEL.JSON beside the written TABLES/."""
    man = manifest(cfg, g, gi, bounds, sp, dtype, scale, chunk_elements, do_escape, esc_iters, esc_tol,
                   escape_scale, robot, ooc, entries, reg_path, field_path, tiling)
    with open(os.path.join(out_dir, "MANIFEST.JSON"), "w") as fh:
        json.dump(man, fh, indent=2)
    mj = model_json(m, reg_path, zeroed, cfg)
    with open(os.path.join(out_dir, "MODEL.JSON"), "w") as fh:
        json.dump(mj, fh, indent=2)
    return man, mj


def manifest(cfg, g, gi, bounds, sp, dtype, scale, chunk_elements, do_escape, esc_iters, esc_tol, escape_scale,
             robot, ooc, entries, reg_path, field_path, tiling=None) -> dict:
    """MANIFEST.JSON, schema 1 (Run.jl:1998-2117, TABLE_FORMAT.md section 2)."""
    p = sp.p
    nbytes, sentinel, _, ebase = card.DTYPES[dtype]
    return {
        "schema_version": 1,
        "generated_utc": datetime.now(timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ"),
        "generator": "peregrine-desktop_amd",
        "regression_sha256": filehash(reg_path),
        "field_sha256": filehash(field_path),
        "grid": {
            "axes": list(AXES), "n": list(g.n), "min": [float(v) for v in g.lo],
            "max": [grid_max(g, k) for k in range(6)],
            "wrap": [False, False, True, False, False, False],
            "units": ["cm", "cm", "rad", "cm/s", "cm/s", "rad/s"],
            "frame": "field",
            "frame_note": "x, y, h and the vx, vy velocity axes are all field frame; omega is frame independent",
            "field_bounds": list(bounds),
            "field_note": "grid.min/max span the positions the robot can legally occupy: the field inset by the "
                          "footprint plus wall_clearance_cm. Clamp x and y into that span as section 3 of "
                          "TABLE_FORMAT.md describes",
            "total_cells": g.ncells,
            "index_formula": "((((ix*Ny+iy)*Nh+ih)*Nvx+ivx)*Nvy+ivy)*Nw+iw",
        },
        "encoding": {
            "dtype": dtype, "elem_bytes": nbytes, "scale": scale, "unit": "seconds",
            "unreachable": sentinel if sentinel is not None else "NaN",
            "escape_base": ebase if do_escape else 0,
            "escape_codes": card.escape_codes(dtype) if do_escape else 0,
            "escape_scale": escape_scale if do_escape else 0.0,
            "escape_unit": "seconds",
            "escape_formula": ("raw >= escape_base -> unreachable, escape_seconds = escape_scale * "
                               "(raw - escape_base)^2") if dtype in ("u8", "u16")
            else "value < 0 -> unreachable, escape_seconds = -value",
            "escape_note": "an escape cell is UNREACHABLE and must lose every comparison against a real route. "
                           "The escape time is only for a robot already in one. See section 6 of TABLE_FORMAT.md",
            "byte_order": "little", "order": "row_major_c",
            "chunk_elements": chunk_elements, "chunk_shift": int(chunk_elements).bit_length() - 1,
        },
        "solver": {
            "dt": float(p.dt), "substeps": p.nsub, "controls": sp.nctl, "control_level": sp.level,
            "iterations_max": sp.iters, "tolerance": sp.tol, "nearest": p.nearest, "control_scan": p.ncoarse,
            "refine_rounds": p.rounds, "refine_delta": p.delta0, "warm_start": sp.warm, "escape": do_escape,
            "escape_iterations": esc_iters, "escape_tolerance": esc_tol, "tau_levels": p.ntau,
            "tau_ratio": p.tau_ratio, "cfl": p.cfl, "tau_min": p.tau_min, "tau_max": p.tau_max,
            "tau_max_auto": sp.tau_auto, "tau_max_value_cost_pct": value_cost(p.tau_max, TAU_REF),
            "substep_max": p.hmax, "adaptive_checks": p.adaptive_checks, "rk2": p.rk2, "simplex": p.simplex,
            "interpolant": "kuhn_simplex_7pt" if p.simplex else "multilinear_64pt",
            "velocity_clamp": p.vclamp, "integrator": "midpoint_rk2" if p.rk2 else "semi_implicit_euler",
            "driver": "tiled" if ooc else "whole_grid", "tiling": tiling,
            "sweep": "jacobi",
            "sweep_checks": p.checks, "zero_c": bool(getc(cfg, "zero_c", True)), "margin_cm": sp.margin,
            "clearance_cm": sp.clearance, "wall_clearance_cm": gi.wall_clearance, "value_cap_s": sp.cap,
            "heading_substeps": gi.hsub, "robot_vertices": 0 if robot is None else int(robot.shape[1]),
            "occupancy": "per heading bin (x, y, h), obstacles and the field boundary alike",
            "resolution": {"xy_cm": float(min(g.step[0], g.step[1])), "heading_deg": math.degrees(float(g.step[2])),
                           "v_cm_s": float(g.step[3]), "w_rad_s": float(g.step[5])},
        },
        "model_file": "MODEL.JSON",
        "targets": entries,
    }


# ---------------------------------------------------------------------------
# CLI (solver/solve.jl)
# ---------------------------------------------------------------------------

USAGE = "usage: solve.py <config.json> [plan|solve|honing [run_dir]]"


def main(argv=None) -> int:
    """solve.jl's modes and exit codes: 0 ok, 1 error (ERROR line and up to
    eight traceback lines on stderr), 2 usage or missing config."""
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv:
        print(USAGE, file=sys.stderr)
        return 2
    cfgpath = argv[0]
    mode = argv[1] if len(argv) >= 2 else "solve"
    if not os.path.isfile(cfgpath):
        print(f"no such config: {cfgpath}", file=sys.stderr)
        return 2
    cfg = _readjson(cfgpath)
    try:
        if mode == "plan":
            print("PLAN " + json.dumps(plan(cfg), separators=(",", ":"), default=_jsonable), flush=True)
        elif mode == "solve":
            run_solve(cfg)
        elif mode == "honing":
            run_dir = argv[2] if len(argv) >= 3 else None
            h = honing_only(cfg, write_dir=run_dir)
            print("HONING " + json.dumps(h, separators=(",", ":")), flush=True)
            if run_dir is not None:
                print("WROTE " + os.path.join(run_dir, "MODEL.JSON"), flush=True)
        else:
            print(f"unknown mode '{mode}' (expected plan, solve or honing)", file=sys.stderr)
            return 2
    except Exception as e:  # noqa: BLE001 -- every failure is reported as ERROR, exit 1
        print(f"ERROR {type(e).__name__}: {e}", file=sys.stderr)
        for line in traceback.format_tb(e.__traceback__)[-8:]:
            print("  " + line.strip().splitlines()[0], file=sys.stderr)
        sys.stderr.flush()
        return 1
    sys.stdout.flush()
    return 0
