"""Host-side tiling arithmetic of `plan` and the solve's backend choice.

Restates, for the report the job interface prints and for the choice between
the in-HBM and the tiled driver:

  * `decompose` (solver/src/Run.jl:316-354): in core or tiled, from
    `out_of_core` (true/false/always/never/auto), `vram_budget_bytes` and
    `vram_headroom`;
  * the reach scan and tile plan the reference sizes its halo with
    (solver/src/Tiles.jl:19-315, 341-411): `horizon_set`, `reach_commands`,
    `reach_extent`, `reach_curve`, `halo_cells`, `plan_tiles`;
  * the step-length advice (solver/src/Advise.jl:19-339, Run.jl:356-404):
    `sweep_work`, the cell cost, `round_seconds`, `tau_options`,
    `recommend_tau`, `fit_advice`, and `settle` (Run.jl:650-682), which
    resolves `tau_max: "auto"`.

The reach scan runs in float64 numpy over the same samples as the reference's
Float32 loop; it sizes advice, not a solve. This library's own tiled driver
does not use the halo: its windows are the exact read ranges of each tile's
boxes (or, for the production candidates, a device reach pass), so a plan's
tile and amplification are the reference's estimate of the trade, kept so
the wizard's reading of PLAN stays meaningful. The cost model is priced with
this build's measured MI355X kernel costs and a host link for the store
(`host_link_rate`, default 25 GB/s of pinned PCIe copies), and with the copies
overlapped with the sweep as the tiled driver runs them (two side streams).
"""
from __future__ import annotations

import math

import numpy as np

from .model import Grid6, Model, Params

TAU_REF = 0.5
# candidate longest steps (Advise.jl:219-220)
TAU_LADDER = (0.5, 0.4, 0.3, 0.25, 0.2, 0.15, 0.12, 0.1, 0.08, 0.06, 0.05, 0.04, 0.03, 0.02)
# Seconds per lookahead and per integration substep or probe on one MI355X,
# from this build's measured sweeps: the plain scheme's LDS-staged sweep of
# problem 2 (1.514 s for 564,854,784 states x 125 lookaheads of 5 Euler
# substeps and 3 probes) split with the reference's per-lookahead share of a
# cell's cost (Advise.jl:109, 0.0765 ns of 0.0765 + 0.1127 (5 + 3) ns).
_LA_S = 1.514 / (564854784 * 125)
_SHARE = 0.07646 / (0.07646 + 0.1127 * 8)
MI355X_CELL_COST = {"per_lookahead_s": _LA_S * _SHARE, "per_step_s": _LA_S * (1 - _SHARE) / 8}
HOST_LINK_RATE = 25e9


def getc(cfg, key, default):
    v = cfg.get(key, default) if isinstance(cfg, dict) else default
    return default if v is None else v


def cell_bytes(warm: bool) -> int:
    """Store bytes per cell: FP32 V and the Int8 warm-start command (Gpu.jl:204-209)."""
    return 7 if warm else 4


# ---------------------------------------------------------------------------
# Horizons and reach (Tiles.jl:19-290)
# ---------------------------------------------------------------------------

def amax_rows(m: Model):
    B = m.B
    return (max(abs(B[0]), abs(B[1]), abs(B[2])), max(abs(B[3]), abs(B[4]), abs(B[5])),
            max(abs(B[6]), abs(B[7]), abs(B[8])))


def cfl_tau(g: Grid6, m: Model, vfx, vfy, w, p: Params):
    """The per-cell horizon (PeregrineSolver.jl:954-976), vectorised over
    speeds and spins, in float64."""
    vfx, vfy, w = np.asarray(vfx, float), np.asarray(vfy, float), np.asarray(w, float)
    shape = np.broadcast(vfx, vfy, w).shape
    if p.cfl <= 0:
        return np.full(shape, float(p.dt))
    ax, ay, aw = amax_rows(m)
    t = np.full(shape, float(p.tau_max))
    for a, st in ((ax, g.step[3]), (ay, g.step[4]), (aw, g.step[5])):
        if a > 1e-6:
            t = np.minimum(t, p.cfl * float(st) / a)
    sp = np.hypot(vfx, vfy)
    lim = p.cfl * min(float(g.step[0]), float(g.step[1])) / np.maximum(sp, 1e-30)
    t = np.where(sp > 1e-3, np.minimum(t, lim), t)
    aw2 = np.abs(w)
    lim = p.cfl * float(g.step[2]) / np.maximum(aw2, 1e-30)
    t = np.where(aw2 > 1e-4, np.minimum(t, lim), t)
    return np.clip(t, p.tau_min, p.tau_max)


def substeps_for(tau, p: Params):
    tau = np.asarray(tau, float)
    if p.hmax <= 0:
        return np.full(tau.shape, int(p.nsub))
    return np.clip(np.trunc(np.minimum(tau / p.hmax, 1e4)).astype(int) + 1, 1, 64)


def horizon_set(g: Grid6, m: Model, p: Params, speed: float, w: float):
    """The horizons cell_update tries at a field speed and spin: tau0, the
    ladder around it and the dt rung (Tiles.jl:19-32)."""
    tau0 = float(cfl_tau(g, m, speed, 0.0, w, p))
    ts = [tau0]
    if p.ntau > 1:
        tau = tau0 * 0.5
        for _ in range(p.ntau):
            ts.append(min(max(tau, p.tau_min), p.tau_max))
            tau *= p.tau_ratio
        if p.cfl > 0:
            ts.append(min(max(p.dt, p.tau_min), p.tau_max))
    return sorted(set(ts))


def reach_commands(ctl: np.ndarray, nsample: int = 512) -> np.ndarray:
    """Coast, the lattice, and a golden-angle spiral of nsample commands
    pushed onto the octahedron surface (Tiles.jl:47-69)."""
    us = [(0.0, 0.0, 0.0)] + [tuple(float(v) for v in row) for row in np.asarray(ctl)]
    ga = math.pi * (3.0 - math.sqrt(5.0))
    for i in range(nsample):
        z = 1.0 - 2.0 * i / max(nsample - 1, 1)
        r = math.sqrt(max(0.0, 1.0 - z * z))
        th = ga * i
        u1, u2, u3 = r * math.cos(th), r * math.sin(th), z
        s = abs(u1) + abs(u2) + abs(u3)
        if s < 1e-6:
            continue
        us.append((u1 / s, u2 / s, u3 / s))
    out = []
    seen = set()
    for u in us:
        if u not in seen:
            seen.add(u)
            out.append(u)
    return np.array(out, float)


def _accel(m: Model, usat, vx, vy, w):
    B, A, S, D = (np.array(x, float).reshape(3, 3) for x in (m.B, m.A, m.S, m.D))
    q, c, eps = np.array(m.q, float), np.array(m.c, float), np.array(m.eps, float)
    v = (vx, vy, w)
    cs = [np.clip(v[k] / eps[k], -1.0, 1.0) for k in range(3)]
    dq = [np.abs(v[k]) * v[k] for k in range(3)]
    out = []
    for r in range(3):
        a = B[r, 0] * usat[0] + B[r, 1] * usat[1] + B[r, 2] * usat[2]
        for k in range(3):
            a = a + A[r, k] * v[k] + S[r, k] * cs[k] + D[r, k] * dq[k]
        out.append(a + q[r] * w * w + c[r])
    return out


def _saturate(m: Model, u):
    if not (m.knee > 0) or math.isnan(m.knee):
        return u
    mag = np.sqrt((u * u).sum(axis=-1, keepdims=True))
    r = mag / m.knee
    gain = np.where(r < 1e-6, 1.0, np.tanh(r) / np.maximum(r, 1e-30))
    return u * gain


def _step(m: Model, p: Params, vx, vy, w, us, tau, nsub):
    """step_state / step_state_rk2 (PeregrineSolver.jl:208-279) from the
    origin at heading 0, vectorised over states and commands."""
    usat = _saturate(m, us)
    usat = [usat[..., 0], usat[..., 1], usat[..., 2]]
    hs = tau / nsub
    x = np.zeros(np.broadcast(vx, usat[0]).shape)
    y = np.zeros_like(x)
    h = np.zeros_like(x)
    vx, vy, w = vx + x, vy + x, w + x

    def deriv(h, vx, vy, w):
        ax, ay, al = _accel(m, usat, vx, vy, w)
        ch, sh = np.cos(h), np.sin(h)
        return ch * vx - sh * vy, sh * vx + ch * vy, w, ax + w * vy, ay - w * vx, al

    with np.errstate(all="ignore"):
        for _ in range(int(nsub)):
            if p.rk2:
                d1 = deriv(h, vx, vy, w)
                mid = [a + b * (0.5 * hs) for a, b in zip((x, y, h, vx, vy, w), d1)]
                d2 = deriv(mid[2], mid[3], mid[4], mid[5])
                x, y, h, vx, vy, w = (a + b * hs for a, b in zip((x, y, h, vx, vy, w), d2))
            else:
                d = deriv(h, vx, vy, w)
                x, y, h, vx, vy, w = (a + b * hs for a, b in zip((x, y, h, vx, vy, w), d))
    return x, y, h, vx, vy, w


def _scan_setup(g: Grid6, nangle: int):
    vhi = [float(g.lo[k]) + float(g.step[k]) * (g.n[k] - 1) for k in range(6)]
    env = [(float(g.lo[k]), vhi[k]) for k in (3, 4, 5)]
    span = math.hypot(float(g.step[0]) * (g.n[0] - 1), float(g.step[1]) * (g.n[1] - 1))
    vmax = max(vhi[3], -float(g.lo[3]), vhi[4], -float(g.lo[4]))
    spmax = math.sqrt(2.0) * vmax
    dsp = 0.5 * min(float(g.step[3]), float(g.step[4]))
    nsp = max(2, math.ceil(spmax / max(dsp, 1e-6)) + 1)
    speeds = np.linspace(0.0, spmax, nsp)
    spins = np.array([float(g.lo[5]) + i * float(g.step[5]) for i in range(g.n[5])])
    th = 2 * math.pi * np.arange(nangle) / nangle
    return env, span, speeds, spins, th


def _reach(m: Model, p: Params, vx, vy, w, us, tau, env, span):
    ns = int(substeps_for(tau, p))
    x, y, h, nvx, nvy, nw = _step(m, p, vx[:, None], vy[:, None], w[:, None], us[None, :, :], tau, ns)
    with np.errstate(all="ignore"):
        ok = np.isfinite(x) & np.isfinite(y) & np.isfinite(h) & np.isfinite(nvx) & np.isfinite(nvy) & np.isfinite(nw)
        d = np.hypot(x, y)
        ok &= d <= span
        if not p.vclamp:
            fx = np.cos(h) * nvx - np.sin(h) * nvy
            fy = np.sin(h) * nvx + np.cos(h) * nvy
            ok &= (fx >= env[0][0]) & (fx <= env[0][1]) & (fy >= env[1][0]) & (fy <= env[1][1])
            ok &= (nw >= env[2][0]) & (nw <= env[2][1])
    dxy = float(np.max(np.where(ok, d, 0.0), initial=0.0))
    dh = float(np.max(np.where(ok, np.abs(h), 0.0), initial=0.0))
    return dxy, dh


def reach_extent(g: Grid6, m: Model, p: Params, ctl: np.ndarray, nangle: int = 48, nsample: int = 512):
    """How far one backup can reach, cm and rad: every horizon of the ladder,
    every command of reach_commands, speeds to the velocity box's corner at
    half a velocity cell, every spin, nangle body-velocity directions; steps
    the backup rejects (diverged, outside the velocity envelope) do not count
    (Tiles.jl:104-187)."""
    us = reach_commands(ctl, nsample)
    env, span, speeds, spins, th = _scan_setup(g, nangle)
    bxy = bh = 0.0
    for sp in speeds:
        vx, vy = sp * np.cos(th), sp * np.sin(th)
        for w in spins:
            wv = np.full_like(vx, w)
            for tau in horizon_set(g, m, p, sp, w):
                a, b = _reach(m, p, vx, vy, wv, us, tau, env, span)
                bxy, bh = max(bxy, a), max(bh, b)
    return bxy, bh


def reach_curve(g: Grid6, m: Model, p: Params, ctl: np.ndarray, taus, nangle: int = 16, nsample: int = 128):
    """The cumulative largest displacement at each horizon of `taus` (sorted):
    entry i is the reach of a ladder capped at taus[i] (Tiles.jl:218-283)."""
    us = reach_commands(ctl, nsample)
    env, span, speeds, spins, th = _scan_setup(g, nangle)
    taus = sorted(set(float(t) for t in taus))
    S, W, T = np.meshgrid(speeds, spins, th, indexing="ij")
    vx, vy, w = (S * np.cos(T)).ravel(), (S * np.sin(T)).ravel(), W.ravel()
    dxy, dh = [], []
    for tau in taus:
        a, b = _reach(m, p, vx, vy, w, us, tau, env, span)
        dxy.append(max(a, dxy[-1] if dxy else 0.0))
        dh.append(max(b, dh[-1] if dh else 0.0))
    return {"taus": taus, "dxy_cm": dxy, "dh_rad": dh}


def reach_at(curve, tau_max: float):
    i = max(0, int(np.searchsorted(curve["taus"], float(tau_max), side="right")) - 1)
    return curve["dxy_cm"][i], curve["dh_rad"][i]


# ---------------------------------------------------------------------------
# Halo and tiles (Tiles.jl:311-411)
# ---------------------------------------------------------------------------

def halo_cells(g: Grid6, dxy_cm: float, margin: int = 2):
    hx = math.ceil(dxy_cm / float(g.step[0])) + 2 + margin
    hy = math.ceil(dxy_cm / float(g.step[1])) + 2 + margin
    return min(hx, g.n[0]), min(hy, g.n[1])


def tile_bytes(g: Grid6, wx: int, wy: int, hx: int, hy: int, warm: bool):
    col = g.ncells // (g.n[0] * g.n[1])
    nxl, nyl = min(wx + 2 * hx, g.n[0]), min(wy + 2 * hy, g.n[1])
    v = nxl * nyl * col * cell_bytes(warm)
    o = nxl * nyl * g.n[2]
    return v + o, nxl, nyl, col


def tile_plan(g: Grid6, wx: int, wy: int, hx: int, hy: int, warm: bool = False, dxy_cm=math.nan,
              dh_rad=math.nan) -> dict:
    b, nxl, nyl, col = tile_bytes(g, wx, wy, hx, hy, warm)
    ntiles = -(-g.n[0] // wx) * -(-g.n[1] // wy)
    return {"hx": hx, "hy": hy, "wx": wx, "wy": wy, "nxl": nxl, "nyl": nyl, "col": col, "bytes": b,
            "ntiles": ntiles, "amplification": (nxl * nyl) / (wx * wy), "dxy_cm": float(dxy_cm),
            "dh_rad": float(dh_rad)}


def plan_tiles(g: Grid6, budget: int, hx: int, hy: int, warm: bool = False, dxy_cm=math.nan, dh_rad=math.nan):
    """The largest tile, interior growing in proportion to the axes, whose
    window fits `budget` bytes; None if not one column does (Tiles.jl:380-400)."""
    n1, n2 = g.n[0], g.n[1]

    def fits(s):
        wx = min(max(round(s * n1), 1), n1)
        wy = min(max(round(s * n2), 1), n2)
        return tile_bytes(g, wx, wy, hx, hy, warm)[0] <= budget, wx, wy

    wx, wy = n1, n2
    if not fits(1.0)[0]:
        if not fits(1.0 / max(n1, n2))[0]:
            return None
        lo, hi = 1.0 / max(n1, n2), 1.0
        for _ in range(48):
            mid = 0.5 * (lo + hi)
            if fits(mid)[0]:
                lo = mid
            else:
                hi = mid
        _, wx, wy = fits(lo)
    return tile_plan(g, wx, wy, hx, hy, warm, dxy_cm, dh_rad)


# ---------------------------------------------------------------------------
# Cost model (Advise.jl:19-210, PeregrineSolver.jl:1028-1081)
# ---------------------------------------------------------------------------

def sweep_work(g: Grid6, m: Model, p: Params, nctl: int) -> dict:
    """Lookaheads, substeps and probes per unblocked cell of one sweep,
    averaged over the velocity sub-grid (PeregrineSolver.jl:1028-1081)."""
    ax, ay, aw = amax_rows(m)
    axy = max(ax, ay)
    dxy = min(float(g.step[0]), float(g.step[1]))
    dh = float(g.step[2])
    nnc = max(nctl - 1, 1)
    span = nnc if p.ncoarse <= 0 else min(p.ncoarse, nnc)
    nfix = 2 + span + 6 * p.rounds
    i4, i5, i6 = np.meshgrid(np.arange(g.n[3]), np.arange(g.n[4]), np.arange(g.n[5]), indexing="ij")
    vfx = float(g.lo[3]) + i4.ravel() * float(g.step[3])
    vfy = float(g.lo[4]) + i5.ravel() * float(g.step[4])
    w = float(g.lo[5]) + i6.ravel() * float(g.step[5])
    speed, spin = np.hypot(vfx, vfy), np.abs(w)
    tau0 = cfl_tau(g, m, vfx, vfy, w, p)

    def nprobes(tau):
        if not p.adaptive_checks:
            return np.full(tau.shape, float(p.checks))
        dpos = speed * tau + 0.5 * axy * tau * tau
        drot = spin * tau + 0.5 * aw * tau * tau
        cells = np.maximum(dpos / (0.5 * dxy), drot / (0.5 * dh))
        return np.maximum(p.checks, np.minimum(np.trunc(np.minimum(cells, 1e4)).astype(int) + 1, 24)).astype(float)

    nL = np.zeros_like(speed)
    nS = np.zeros_like(speed)
    nP = np.zeros_like(speed)

    def add(t, mult):
        nonlocal nL, nS, nP
        nL = nL + mult
        nS = nS + mult * substeps_for(t, p)
        nP = nP + mult * nprobes(t)

    add(tau0, nfix)
    if p.ntau > 1:
        tau = tau0 * 0.5
        for k in range(1, p.ntau + 1):
            if k != 2:
                add(np.clip(tau, p.tau_min, p.tau_max), 1)
            tau = tau * p.tau_ratio
        if p.cfl > 0:
            add(np.full_like(tau0, min(max(p.dt, p.tau_min), p.tau_max)), 1)
    return {"lookaheads": float(nL.mean()), "substeps": float(nS.mean()), "probes": float(nP.mean())}


def cell_cost(cfg) -> tuple:
    """(per_lookahead_s, per_step_s), source: `cell_cost` from the config
    when it is there (Advise.jl:83-106), else this build's measured MI355X
    costs."""
    cc = getc(cfg, "cell_cost", None)
    if isinstance(cc, dict) and float(cc.get("per_step_s", 0.0)) > 0:
        return {"per_lookahead_s": float(cc.get("per_lookahead_s", MI355X_CELL_COST["per_lookahead_s"])),
                "per_step_s": float(cc["per_step_s"])}, "cost"
    return dict(MI355X_CELL_COST), "measured"


def cell_seconds(work: dict, cost: dict) -> float:
    return cost["per_lookahead_s"] * work["lookaheads"] + cost["per_step_s"] * (work["substeps"] + work["probes"])


def effective_rate(work: dict, cost: dict) -> float:
    return 1.0 / max(cell_seconds(work, cost), 1e-18)


def round_seconds(cells: float, sweeps: int, amp: float, sb: float, rate: float, link: float) -> dict:
    """One round: sweeps over every cell, and the window loads and interior
    stores over the host link. The tiled driver copies on two side streams
    while a tile sweeps, so the round is the larger of the two (the reference's
    disk-bound driver sums them, Advise.jl:168-210)."""
    compute = cells * sweeps / rate
    io = cells * sb * amp / link + cells * sb / link
    return {"total": max(compute, io), "compute": compute, "io": io}


def value_cost(tau: float, frm: float | None = None) -> float:
    from .run import value_cost as vc
    return vc(tau, frm)


def retau(p: Params, tau: float) -> Params:
    from dataclasses import replace
    return replace(p, tau_max=float(tau))


def tau_options(g: Grid6, m: Model, p: Params, ctl, nctl: int, budget: int, warm: bool, margin: int, nang: int,
                ncmd: int, sb: float, sweeps: int, cost: dict, link: float, ladder=TAU_LADDER) -> list:
    """Every candidate longest step with a workable tiling, longest first, with
    its reach, halo, tile, amplification, round time split and value cost
    (Advise.jl:252-273)."""
    taus = sorted(set(float(t) for t in ladder))
    curve = reach_curve(g, m, p, ctl, taus, nangle=nang, nsample=ncmd)
    cells = float(g.ncells)
    # the sweeps a round costs at the reference horizon: what the same grid
    # would cost held whole, against which only the copies are a tiling cost
    ref_compute = cells * sweeps / effective_rate(sweep_work(g, m, retau(p, TAU_REF), nctl), cost)
    opts = []
    for tau in reversed(taus):
        dxy, dh = reach_at(curve, tau)
        hx, hy = halo_cells(g, dxy, margin)
        tp = plan_tiles(g, budget, hx, hy, warm, dxy, dh)
        if tp is None:
            continue
        rate = effective_rate(sweep_work(g, m, retau(p, tau), nctl), cost)
        r = round_seconds(cells, sweeps, tp["amplification"], sb, rate, link)
        opts.append({"tau": tau, "reach_cm": dxy, "halo": hx, "tile": tp["wx"], "amplification": tp["amplification"],
                     "round_s": r["total"], "compute_s": r["compute"], "io_s": r["io"], "cell_rate": rate,
                     "tiling_s": max(ref_compute, r["io"]), "cost": value_cost(tau, TAU_REF)})
    return opts


def recommend_tau(mode: str, g: Grid6, m: Model, p: Params, ctl, nctl: int, budget: int, warm: bool, margin: int,
                  nang: int, ncmd: int, sb: float, sweeps: int, cost: dict, link: float) -> dict:
    """The step minimising the tiled round's time x (1 + value cost), longest
    first on ties; in core the reference horizon with no scan
    (Advise.jl:304-339)."""
    cur = float(p.tau_max)
    if mode != "ooc":
        return {"tau_max": TAU_REF, "cost": 0.0, "current": cur, "current_cost": value_cost(cur, TAU_REF),
                "io_minor": True, "options": [],
                "reason": "the grid is held whole in HBM, so the lookahead sets no halo and costs nothing but a few "
                          "integration substeps"}
    opts = tau_options(g, m, p, ctl, nctl, budget, warm, margin, nang, ncmd, sb, sweeps, cost, link)
    if not opts:
        return {"tau_max": None, "cost": None, "current": cur, "current_cost": None, "io_minor": False,
                "options": opts, "reason": "no tiling of this grid fits the device budget at any step length"}
    # A shorter step also integrates fewer substeps, but a grid held whole
    # would not shorten it for that; what tiling adds is the copies, so the
    # time scored is a round at the reference horizon's compute, or the
    # copies when they cannot hide behind it (the reference scores its own
    # disk-bound round, Advise.jl:326).
    best = min(opts, key=lambda o: o["tiling_s"] * (1.0 + o["cost"] / 100))
    minor = best["io_s"] <= 0.25 * best["compute_s"]
    return {"tau_max": best["tau"], "cost": best["cost"], "current": cur, "current_cost": value_cost(cur, TAU_REF),
            "io_minor": minor, "options": opts,
            "reason": ("the window copies are a small part of a round here, so a shorter step would trade "
                       "accuracy for little time") if minor else
                      ("where the trade turns: a shorter step saves less time than the accuracy it costs, a "
                       "longer one costs more time than the accuracy is worth")}


def fit_advice(cfg, g: Grid6, m: Model, p: Params, ctl, nctl: int, budget: int, warm: bool, margin: int, nang: int,
               ncmd: int) -> dict:
    """When no tiling fits: the largest halo the budget admits, whether
    dropping the warm start alone would fit, and the longest steps that fit
    (the first, and the first loading at most 8 cells per cell updated)
    (Run.jl:356-404)."""
    col = g.ncells // (g.n[0] * g.n[1])
    per = cell_bytes(warm)
    maxcols = budget // (col * per)
    hmax = -1 if maxcols <= 0 else (math.isqrt(maxcols) - 1) // 2
    cost, _ = cell_cost(cfg)
    link = float(getc(cfg, "host_link_rate", HOST_LINK_RATE))
    opts = [o for o in tau_options(g, m, p, ctl, nctl, budget, warm, margin, nang, ncmd, float(per), 1, cost, link)
            if o["tau"] < float(p.tau_max)]
    comfy = next((o for o in opts if o["amplification"] <= 8.0), None)
    warm_helps = False
    if warm:
        dxy, dh = reach_extent(g, m, p, ctl, nang, ncmd)
        hx, hy = halo_cells(g, dxy, margin)
        warm_helps = plan_tiles(g, budget, hx, hy, False, dxy, dh) is not None
    return {"column_bytes": col * per, "budget_bytes": budget, "max_halo": hmax, "warm_helps": warm_helps,
            "tau_max": opts[0] if opts else None, "tau_comfortable": comfy}


def out_of_core_choice(cfg):
    """(force_on, force_off) from `out_of_core` (Run.jl:317-319)."""
    want = getc(cfg, "out_of_core", "auto")
    on = want is True or want in ("true", "always")
    off = want is False or want in ("false", "never")
    return on, off


def decompose(cfg, g: Grid6, m: Model, p: Params, ctl, nctl: int, warm: bool, device_budget) -> dict:
    """In core or tiled (Run.jl:316-354). `device_budget(headroom)` gives the
    device bytes a run may use (None without a device). The budget is
    `vram_budget_bytes` when positive, else the device's free bytes times
    `vram_headroom`; the whole grid needs the Jacobi pair (and the warm-start
    store) on the device. `out_of_core` forces the choice either way."""
    force_on, force_off = out_of_core_choice(cfg)
    pinned = int(getc(cfg, "vram_budget_bytes", 0))
    budget = pinned
    if budget <= 0:
        b = device_budget(float(getc(cfg, "vram_headroom", 0.85)))
        budget = int(b) if b is not None else 0
    whole = g.ncells * (8 + (3 if warm else 0))
    col = g.ncells // (g.n[0] * g.n[1])
    out = {"budget": budget, "whole_bytes": whole, "col": col, "pinned_budget": pinned > 0}
    if budget <= 0:
        return {**out, "mode": "cpu", "tp": None}
    if force_off or (not force_on and whole <= budget):
        return {**out, "mode": "incore", "tp": None}
    nang = int(getc(cfg, "halo_scan_angles", 16))
    ncmd = int(getc(cfg, "halo_scan_commands", 128))
    margin = int(getc(cfg, "halo_margin", 2))
    wt = bool(getc(cfg, "warm_start_tiled", True))
    dxy, dh = reach_extent(g, m, p, ctl, nang, ncmd)
    hx, hy = halo_cells(g, dxy, margin)
    tp = plan_tiles(g, budget, hx, hy, wt, dxy, dh)
    out.update(mode="ooc", tp=tp, dxy_cm=dxy, dh_rad=dh, hx=hx, hy=hy, advice=None)
    if tp is None:
        out["advice"] = fit_advice(cfg, g, m, p, ctl, nctl, budget, wt, margin, nang, ncmd)
    return out


def settle(cfg, g: Grid6, m: Model, p: Params, ctl, nctl: int, warm: bool, tau_auto: bool, device_budget) -> dict:
    """Resolve `tau_max: "auto"` (Run.jl:650-682): in core the reference
    horizon; tiled, recommend_tau's pick (never longer than the reference).
    Returns the parameters in force, the decomposition and the advice."""
    cost, _ = cell_cost(cfg)
    link = float(getc(cfg, "host_link_rate", HOST_LINK_RATE))
    wt = bool(getc(cfg, "warm_start_tiled", True))
    margin = int(getc(cfg, "halo_margin", 2))
    nang = int(getc(cfg, "halo_scan_angles", 16))
    ncmd = int(getc(cfg, "halo_scan_commands", 128))
    if not tau_auto:
        return {"p": p, "dec": decompose(cfg, g, m, p, ctl, nctl, warm, device_budget), "auto": False, "rec": None}
    p0 = retau(p, TAU_REF)
    dec0 = decompose(cfg, g, m, p0, ctl, nctl, warm, device_budget)
    rec = recommend_tau(dec0["mode"], g, m, p0, ctl, nctl, dec0["budget"], wt, margin, nang, ncmd,
                        float(cell_bytes(wt)), 1, cost, link)
    if rec["tau_max"] is None or rec["tau_max"] >= TAU_REF:
        return {"p": p0, "dec": dec0, "auto": True, "rec": rec}
    p1 = retau(p, rec["tau_max"])
    return {"p": p1, "dec": decompose(cfg, g, m, p1, ctl, nctl, warm, device_budget), "auto": True, "rec": rec}
