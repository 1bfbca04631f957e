"""The precision form's accuracy check at full problem-2 size (564,854,784
states, 125 controls, six obstacles), run by hand on a GPU host in two
steps (each fits one run of about 15 minutes):

    python3 tests/full_size_accuracy.py truth  SAMPLE   # FP64 table, certified
    python3 tests/full_size_accuracy.py compare SAMPLE  # both FP32 forms vs it

truth: the device's diagonal-solved FP64 Jacobi table from cap to
||dV||inf < 1e-6 s; 20,000 reached, non-target states of a seeded
1,000,000-state sample (numpy PCG64(3 + 100 * 2), uniform over the grid) are
backed up by the CPU FP64 oracle against the whole downloaded table, and
max |T V - V| over them is recorded beside the sample (indices, FP64 values
and FP64 argmins) in SAMPLE.

compare: the exact FP32 table (the reference's arithmetic) to the first sweep
with ||dV||inf < 1e-3 s and then until a sweep changes nothing; the
precision form (16-bit store) to its own ||dV||inf < 1e-3 s and then exact
FP32 polish sweeps until one moves no state by 1e-3 s or more. Over the
sampled states both reach: errors against FP64 (max, p50, p99, p99.9), the
states where the stagnated exact FP32 table is more than 5e-3 s from FP64
(a discrete decision FP32 takes the other way; excluded and listed), the
precision form's maximum outside them, and argmin agreement. Prints one
JSON line.
"""
from __future__ import annotations

import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from conftest import load_package, oracle_problem  # noqa: E402

CONFIG = 2
TOL = 1e-3
BAR = 5e-3


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def sample_indices(n):
    rng = np.random.Generator(np.random.PCG64(3 + 100 * CONFIG))
    return np.sort(rng.integers(0, n, 1_000_000))


def truth(pg, prob, out):
    from oracle import oracle as orc
    orc.lib()
    g = prob.grid
    cap = prob.params.cap
    t = time.perf_counter()
    with pg.HipSolver(g, prob.model, prob.params, prob.ctl, prob.occ) as s:
        s.fill_seed(prob.seeds)
        s.solve_f64(0, 0.0)
        its, l = 0, float("inf")
        for its in range(1, 2001):
            l, _ = s.sweep_f64(diagonal=True)
            if its % 10 == 0:
                log(f"FP64 sweep {its}: linf {l:.3e}, {time.perf_counter() - t:.1f} s")
            if l < 1e-6:
                break
        r64, _ = s.residual_f64(0)
        V64 = s.download_f64()
        P64 = s.policy_f64()
    t64 = time.perf_counter() - t
    idx = sample_indices(g.ncells)
    seeds = np.zeros(g.ncells, bool)
    seeds[prob.seeds] = True
    reached = idx[(V64[idx] < cap) & ~seeds[idx]]
    rng = np.random.default_rng(11)
    cells = np.sort(rng.choice(reached, 20_000, replace=False))
    op = oracle_problem(orc, prob)
    t = time.perf_counter()
    TV, _ = op.backup_cells_f64(V64, cells, nthreads=min(16, os.cpu_count() or 1))
    cert = float(np.max(np.abs(TV - V64[cells])))
    log(f"oracle FP64 backups of {cells.size} reached sampled states: max |T V - V| {cert:.2e} "
        f"({time.perf_counter() - t:.1f} s)")
    info = {"sweeps": its, "linf": l, "seconds": t64, "device_residual_linf": r64,
            "mode": "diagonal-solved FP64 Jacobi from cap", "oracle_certified_states": int(cells.size),
            "oracle_max_residual": cert,
            "sample": "1,000,000 states, numpy PCG64(3 + 100 * config), uniform over the grid"}
    log(f"FP64 table: {json.dumps(info)}")
    with open(out, "wb") as f:
        np.savez(f, idx=idx, v64=V64[idx], p64=P64[idx], info=json.dumps(info), config=CONFIG)
    print(json.dumps(info), flush=True)


def solve(pg, prob, vstore, tag, stagnate=False):
    g = prob.grid
    s = pg.HipSolver(g, prob.model, prob.params, prob.ctl, prob.occ, vstore=vstore)
    s.fill_seed(prob.seeds)
    t = time.perf_counter()
    its = 0
    for its in range(1, 3001):
        r = s.sweep(its - 1, its % 2 == 1)
        if its % 10 == 0:
            log(f"{tag} sweep {its}: linf {r.linf:.5f}, {time.perf_counter() - t:.1f} s")
        if r.linf < TOL:
            break
    res = {"sweeps": its, "seconds": time.perf_counter() - t}
    if vstore != pg.abi.PG_VSTORE_F32:
        s.polish(0)
        trace = []
        for k in range(200):
            r = s.sweep(k, k % 2 == 1)
            trace.append(r.linf)
            log(f"{tag} polish sweep {k}: exact linf {r.linf:.5f}")
            if r.linf < TOL:
                break
        res["polish_linf"] = trace
        res["seconds_with_polish"] = time.perf_counter() - t
    V = s.download()
    P = s.policy()
    Vc = None
    if stagnate:
        k = 0
        for k in range(1, 3001):
            r = s.sweep(its + k - 1, (its + k) % 2 == 1)
            if k % 10 == 0:
                log(f"{tag} to stagnation, sweep {k}: linf {r.linf:.2e}")
            if r.linf == 0.0:
                break
        res["sweeps_to_stagnation"] = its + k
        Vc = s.download()
    s.close()
    log(f"{tag}: {json.dumps(res)}")
    return res, V, P, Vc


def stats(e):
    q = np.quantile(e, [0.5, 0.99, 0.999])
    return {"max": float(e.max()), "p50": float(q[0]), "p99": float(q[1]), "p999": float(q[2])}


def compare(pg, prob, sample):
    z = np.load(sample, allow_pickle=False)
    if int(z["config"]) != CONFIG:
        raise SystemExit(f"{sample} holds problem {int(z['config'])}")
    idx, V64, P64 = z["idx"], z["v64"], z["p64"]
    info = json.loads(str(z["info"]))
    g = prob.grid
    cap = prob.params.cap
    r32, Ve, Pe, Vc = solve(pg, prob, pg.abi.PG_VSTORE_F32, "fp32", stagnate=True)
    r16, Vp, Pp, _ = solve(pg, prob, pg.abi.PG_VSTORE_U16_TILE, "u16_tile")
    ve, vc, vp = (a[idx].astype(np.float64) for a in (Ve, Vc, Vp))
    both = (ve < cap) & (vc < cap) & (vp < cap) & (V64 < cap)
    ee, ec, ep = np.abs(ve - V64), np.abs(vc - V64), np.abs(vp - V64)
    off = both & (ec > BAR)
    keep = both & ~off
    seeds = np.zeros(g.ncells, bool)
    seeds[prob.seeds] = True
    blocked = np.asarray(prob.occ, np.uint8).reshape(-1)[idx // (g.n[3] * g.n[4] * g.n[5])] != 0
    live = both & ~blocked & ~seeds[idx]
    out = {
        "f64": info, "fp32_run": r32, "u16_tile_run": r16, "n_both_reached": int(both.sum()),
        "reach_mismatch_u16_vs_f64": float(np.mean((vp < cap) != (V64 < cap))),
        "fp32_at_1e-3": {**stats(ee[both]), "max_outside_excluded": float(ee[keep].max())},
        "fp32_stagnated": stats(ec[both]),
        "u16_tile_polished": {**stats(ep[both]), "max_outside_excluded": float(ep[keep].max()),
                              "above_bar_outside_excluded": int(np.count_nonzero(ep[keep] > BAR))},
        "excluded_states": int(off.sum()),
        "excluded_sample": [(int(idx[i]), round(float(vc[i] - V64[i]), 4), round(float(vp[i] - V64[i]), 4))
                            for i in np.nonzero(off)[0][:12]],
        "argmin_agreement_fp32": float(np.mean(Pe[idx][live] == P64[live])),
        "argmin_agreement_u16_tile": float(np.mean(Pp[idx][live] == P64[live])),
    }
    print(json.dumps(out), flush=True)


def main():
    pg = load_package()
    prob = pg.configs.problem(CONFIG)
    mode, path = sys.argv[1], sys.argv[2]
    if mode == "truth":
        truth(pg, prob, path)
    elif mode == "compare":
        compare(pg, prob, path)
    else:
        raise SystemExit("usage: full_size_accuracy.py truth|compare SAMPLE")


if __name__ == "__main__":
    main()
