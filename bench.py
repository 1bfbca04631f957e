"""Benchmark: state-updates/s and time to ||dV||inf < 1e-3 s of the plain
value-iteration solve on one MI355X, FP32 and the 16-bit precision form.

A step is one Bellman sweep over the whole grid (every state backed up once
against every control), inputs resident in HBM. The default workload is the
564,854,784-state, 125-control problem (64x64x48x13x13x17 grid, six
obstacles, plain FP32 scheme). The default run prints one JSON line with:

  value          FP32 sweep throughput (every box computed), state-updates/s
  converge       time from V = cap to the first sweep with ||dV||inf < 1e-3 s,
                 for FP32 and for the precision form (16-bit sweeps to its own
                 1e-3, then exact FP32 polish sweeps until an exact sweep moves
                 no cell by 1e-3 s or more), and the speed-up
  u16_tile       the precision form's sweep throughput (16-bit fixed-point
                 codes per 4^6-cell tile, FP32 own values, contracted FP32
                 backups with the diagonal solved)
  fp32_contracted  the FP32 store swept by the precision form's contracted
                 arithmetic: separates what the 16-bit store buys from what
                 the arithmetic buys
  roofline       the FP32 sweep kernel against its binding resource, vector
                 ALU issue: the reference algorithm's floating-point operations
                 per state-update over the kernel time, against the FP32
                 vector peak; beside it the kernel's measured HBM traffic and
                 the metric's logical gather bytes (LDS-served, not HBM)
  cpu_baseline   the oracle's in-place threaded sweep (the reference's
                 sweep_cpu! order) over a contiguous slab of the same grid

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2]
                    [--no-converge] [--accuracy] [--no-cpu] [--cpu-converge]

--accuracy also solves the FP64 table of the same problem and reports both
forms' errors against it on a seeded 1M-state subsample, with argmin-policy
agreement (under "accuracy"). --f64-only solves just the FP64 table and writes
the subsample (indices, FP64 values, FP64 policy) to --f64-out; --f64-sample
reads such a file in place of the FP64 solve. --cpu-converge times the oracle
from V = cap to ||dV||inf < 1e-3 on problem 1 in place (Gauss-Seidel,
alternating direction) and Jacobi.

--adjoint computes dL/dtheta for the 30 drivetrain parameters (default
workload: problem 5, L = mean V over 4096 seeded starts) on the FP64 table and
checks --fd-params of them against central differences of L.

--ooc runs the out-of-core sweep instead (default workload: problem 4,
96x96x64x17x17x21, device budget 8 GB, 16x16x8 tiles): --steps tiled sweeps
from the seeded start, timed, each compared bit for bit with the in-HBM sweep
of the same grid (--ooc-compare, on by default when it fits).
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0
FP32_VECTOR_PEAK_TFLOPS = 157.3
# Counter figures per launch of a sweep kernel, keyed by (problem, kernel
# id), from rocprofv3 --pmc passes over this script (one sweep per leg):
#   hbm_bytes      FETCH_SIZE + WRITE_SIZE, each in a pass of its own, in
#                  KiB x 1024 (scale factor 1.0: WRITE_SIZE reads 1,048,576 KiB
#                  for a 1 GiB device fill)
#   flops_counted  64 x (SQ_INSTS_VALU_FLOPS_FP32 + _FP64 + their _TRANS
#                  counterparts): the counters sum each wave instruction's
#                  per-lane FLOPs, so 64 lanes make the executed FLOPs
#   valu_insts     SQ_INSTS_VALU, wave64 VALU instructions issued
#   valu_issue_frac  valu_insts over the issue slots of the launch: 1024 SIMDs,
#                  one wave64 FP32 instruction per SIMD per 2 cycles (the rate
#                  behind the 157.3 TFLOP/s FP32 peak), for GRBM_GUI_ACTIVE / 8
#                  cycles (GRBM_GUI_ACTIVE sums the 8 XCDs)
#   lds_bank_conflict_cycles  SQ_LDS_BANK_CONFLICT
# Counters cannot be read inside the timed run, so the figures are carried.
# Kernels: 2 = exact LDS-staged sweep (the FP32 table), 3 = contracted sweep
# over the FP32 table, 4 = contracted sweep over the 16-bit store.
PMC = {
    (2, 2): {"hbm_bytes": (21_171_403 + 2_660_072) * 1024, "flops_counted": 64 * 685_315_275_294,
             "valu_insts": 1_011_039_627_850, "valu_issue_frac": 0.701,
             "lds_bank_conflict_cycles": 195_850_124_160},
    (2, 3): {"hbm_bytes": (21_111_973 + 2_717_123) * 1024, "flops_counted": 64 * 467_560_510_574,
             "valu_insts": 479_103_616_971, "valu_issue_frac": 0.634,
             "lds_bank_conflict_cycles": 195_867_540_288},
    (2, 4): {"hbm_bytes": (8_573_177 + 2_591_681) * 1024, "flops_counted": 64 * 486_117_086_529,
             "valu_insts": 541_009_196_454, "valu_issue_frac": 0.661,
             "lds_bank_conflict_cycles": 195_867_540_288},
}


def load_package():
    name = "peregrine_desktop_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg = ROOT / "peregrine-desktop_amd"
    spec = importlib.util.spec_from_file_location(name, pkg / "__init__.py",
                                                  submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def host_threads() -> int:
    """The host threads the CPU baseline may use: OMP_NUM_THREADS when the
    environment sets it (the GPU host's share), else every core."""
    v = os.environ.get("OMP_NUM_THREADS", "")
    return int(v) if v.isdigit() and int(v) > 0 else (os.cpu_count() or 1)


def cpu_baseline(pg, prob, V, seconds_target: float = 15.0):
    """The oracle's in-place, OpenMP-threaded sweep -- the reference CPU
    path's update order (sweep_cpu!, solver/src/Pipeline.jl:552-568) -- over
    a contiguous slab of cells in the middle of the grid, from the V the GPU
    warm-up sweeps left. A full problem-2 sweep takes the host tens of
    minutes, so the slab is sized to ~seconds_target. Rank 0 only."""
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc

    nthreads = host_threads()
    op = orc.Problem(prob.grid.c(), prob.model.c_struct(), prob.params.c(), pg.ctl_soa(prob.ctl),
                     prob.occ)
    V = np.ascontiguousarray(V, dtype=np.float32).copy()
    mid = V.size // 2
    n = 32768
    t = time.perf_counter()
    op.sweep_range(V, mid - n // 2, mid + n // 2, rev=0, phase=0, nthreads=nthreads)
    dt = time.perf_counter() - t
    n = int(min(V.size, max(n, n * seconds_target / max(dt, 1e-6))))
    lo = max(0, mid - n // 2)
    t = time.perf_counter()
    op.sweep_range(V, lo, lo + n, rev=1, phase=1, nthreads=nthreads)
    dt = time.perf_counter() - t
    ix0 = lo // (V.size // prob.grid.n[0])
    return {"value": n / dt, "unit": "state-updates/s", "cores": nthreads, "kind": "port",
            "sample": f"in-place threaded sweep (reference CPU order, backward pass) of {n} contiguous cells "
                      f"from flat index {lo} (x plane {ix0}) of the same grid, {prob.nctl}-control backups, "
                      f"against the V after the GPU warm-up sweeps; {dt:.1f} s"}


def cpu_converge(pg, args) -> dict:
    """Problem 1 (537,600 states, 27 controls) on the host from V = cap to
    the first sweep with ||dV||inf < tol: in place with alternating direction
    (Gauss-Seidel, the reference's CPU order) and Jacobi, OpenMP over cells.
    Per-sweep time excludes the first sweep (Run.jl:1835-1848)."""
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc

    prob = pg.configs.problem(1)
    op = orc.Problem(prob.grid.c(), prob.model.c_struct(), prob.params.c(), pg.ctl_soa(prob.ctl),
                     prob.occ)
    nthreads = host_threads()
    out = {"workload": "problem 1: 16x16x12x5x5x7, 27 controls", "cores": nthreads, "kind": "port"}
    for tag, mode in (("gauss_seidel", 2), ("jacobi", 0)):
        V = np.full(prob.grid.ncells, prob.params.cap, np.float32)
        V[prob.seeds] = 0
        W = np.empty_like(V)
        per = []
        t0 = time.perf_counter()
        for it in range(1, 5001):
            t = time.perf_counter()
            if mode == 0:
                _, l = op.sweep(V, W, mode=0, phase=it - 1, nthreads=nthreads)
                V, W = W, V
            else:
                _, l = op.sweep(V, None, mode=2, rev=it % 2 == 1, phase=it - 1, nthreads=nthreads)
            V[prob.seeds] = 0
            per.append(time.perf_counter() - t)
            if it % 50 == 0:
                print(f"# cpu {tag} sweep {it}: linf {l:.5f}, {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                      flush=True)
            if l < args.tol:
                break
        out[tag] = {"sweeps": it, "seconds": time.perf_counter() - t0, "linf": l,
                    "seconds_per_sweep": float(np.mean(per[1:])) if len(per) > 1 else per[0],
                    "state_updates_per_s": prob.grid.ncells / (float(np.mean(per[1:])) if len(per) > 1 else per[0])}
        print(f"# cpu {tag}: {json.dumps(out[tag])}", file=sys.stderr, flush=True)
    return out


def solve_timed(pg, prob, vstore, args, tag):
    """From V = cap (targets 0) to the first sweep with ||dV||_inf < tol;
    wall time of the sweeps (host loop included), the sweep count, and the
    table and policy at the end. A 16-bit store is then polished: exact
    FP32 sweeps until one moves no cell by tol or more (at most
    --polish-max), so its time to tolerance is that of the exact backup."""
    g = prob.grid
    s = pg.HipSolver(g, prob.model, prob.params, prob.ctl, prob.occ, vstore=vstore)
    s.set_active(not args.no_active)
    s.fill_seed(prob.seeds)
    t = time.perf_counter()
    its, linf = 0, float("inf")
    computed = total = 0
    for its in range(1, args.max_sweeps + 1):
        r = s.sweep(its - 1, its % 2 == 1)
        linf = r.linf
        c, nb = s.last_sweep_boxes()
        computed += c
        total += nb
        if its % 10 == 0:
            print(f"# converge {tag} sweep {its}: linf {linf:.5f}, {time.perf_counter() - t:.1f} s, "
                  f"boxes computed {c / nb:.3f}", file=sys.stderr, flush=True)
        if linf < args.tol:
            break
    t_conv = time.perf_counter() - t
    res = {"seconds": t_conv, "sweeps": its, "linf": linf, "tol": args.tol, "converged": bool(linf < args.tol),
           "active_boxes": not args.no_active, "boxes_computed_frac": computed / max(total, 1)}
    if vstore != pg.abi.PG_VSTORE_F32:
        t = time.perf_counter()
        s.polish(0)
        lp, trace = float("inf"), []
        for k in range(args.polish_max):
            r = s.sweep(k, k % 2 == 1)
            lp = r.linf
            trace.append(lp)
            print(f"# converge {tag} polish sweep {k}: exact linf {lp:.5f}", file=sys.stderr, flush=True)
            if lp < args.tol:
                break
        res["polish_sweeps"] = len(trace)
        res["polish_seconds"] = time.perf_counter() - t
        res["polish_linf"] = trace
        res["polish_converged"] = bool(lp < args.tol)
        res["seconds_with_polish"] = t_conv + res["polish_seconds"]
    V = s.download()
    pol = s.policy()
    res["reached_frac"] = float(np.count_nonzero(V < prob.params.cap) / V.size)
    print(f"# converge {tag}: {json.dumps(res)}", file=sys.stderr, flush=True)
    s.close()
    return res, V, pol


def converge_both(pg, prob, args):
    r32, V32, p32 = solve_timed(pg, prob, pg.abi.PG_VSTORE_F32, args, "fp32")
    r16, V16, p16 = solve_timed(pg, prob, pg.abi.PG_VSTORE_U16_TILE, args, "u16_tile")
    r16["speedup_vs_fp32"] = r32["seconds"] / r16["seconds_with_polish"]
    return {"fp32": r32, "u16_tile": r16}, {"fp32": (V32, p32), "u16_tile": (V16, p16)}


def f64_sample(pg, prob, args):
    """The FP64 table of the problem (diagonal-solved FP64 Jacobi from cap to
    ||dV||_inf < f64_tol) at a seeded 1M-state subsample: numpy
    PCG64(3 + 100 * config), uniform over the grid."""
    g = prob.grid
    s = pg.HipSolver(g, prob.model, prob.params, prob.ctl, prob.occ)
    s.fill_seed(prob.seeds)
    t = time.perf_counter()
    s.solve_f64(0, 0.0)
    its, l = 0, float("inf")
    for its in range(1, args.f64_max_sweeps + 1):
        l, _ = s.sweep_f64(diagonal=True)
        if its % 10 == 0:
            print(f"# FP64 sweep {its}: linf {l:.3e}, {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
        if l < args.f64_tol:
            break
    t64 = time.perf_counter() - t
    r64, _ = s.residual_f64(0)
    rng = np.random.Generator(np.random.PCG64(3 + 100 * args.config))
    idx = np.sort(rng.integers(0, g.ncells, 1_000_000))
    V64 = s.download_f64()[idx]
    P64 = s.policy_f64()[idx]
    s.close()
    info = {"sweeps": its, "linf": l, "tol": args.f64_tol, "seconds": t64, "residual_linf": r64,
            "mode": "diagonal-solved FP64 Jacobi from cap",
            "sample": "1,000,000 states, numpy PCG64(3 + 100 * config), uniform over the grid"}
    print(f"# FP64 table: {json.dumps(info)}", file=sys.stderr, flush=True)
    return idx, V64, P64, info


def accuracy(pg, prob, args, tables):
    """Both forms' tables against the FP64 table of the same problem on the
    seeded 1M-state subsample; errors where both tables are below cap,
    policy agreement over those states that are not blocked."""
    g = prob.grid
    cap = prob.params.cap
    if args.f64_sample:
        z = np.load(args.f64_sample, allow_pickle=False)
        idx, V64, P64 = z["idx"], z["v64"], z["p64"]
        info = json.loads(str(z["info"]))
        if int(z["config"]) != args.config:
            raise SystemExit(f"{args.f64_sample} holds problem {int(z['config'])}, not {args.config}")
    else:
        idx, V64, P64, info = f64_sample(pg, prob, args)
    out = {"f64": info}
    blocked = np.asarray(prob.occ, np.uint8).reshape(-1)[idx // (g.n[3] * g.n[4] * g.n[5])] != 0
    for tag, (V, P) in tables.items():
        a, b = V[idx].astype(np.float64), V64
        both = (a < cap) & (b < cap)
        e = np.abs(a - b)[both]
        q = np.quantile(e, [0.5, 0.99, 0.999])
        live = both & ~blocked
        sd = (a - b)[both]
        jm = int(np.argmax(e))
        out[tag] = {"max": float(e.max()), "p50": float(q[0]), "p99": float(q[1]), "p999": float(q[2]),
                    "max_cell": int(idx[both][jm]), "max_signed": float(sd[jm]),
                    "above_1e-3": int(np.count_nonzero(e > 1e-3)), "above_5e-3": int(np.count_nonzero(e > 5e-3)),
                    "mean_signed": float(np.mean((a - b)[both])), "n": int(both.sum()),
                    "reach_mismatch": float(np.mean((a < cap) != (b < cap))),
                    "policy_agreement": float(np.mean(P[idx][live] == P64[live]))}
        print(f"# accuracy {tag}: {json.dumps(out[tag])}", file=sys.stderr, flush=True)
    return out


def run_adjoint(pg, prob, args) -> dict:
    """dL/dtheta by the adjoint against central differences of L."""
    g = prob.grid
    starts = pg.configs.adjoint_starts(prob, 4096)

    def solve(model):
        t = time.perf_counter()
        with pg.HipSolver(g, model, prob.params, prob.ctl, prob.occ) as s:
            s.fill_seed(prob.seeds)
            its, l = s.solve_f64(50000, args.f64_tol)
            t_solve = time.perf_counter() - t
            t = time.perf_counter()
            L, grad, it = s.adjoint(starts)
            return L, grad, its, l, t_solve, time.perf_counter() - t, it

    L, grad, its, l, t_solve, t_adj, it_adj = solve(prob.model)
    print(f"# adjoint: L {L:.6f}, FP64 table {its} sweeps ({t_solve:.1f} s, linf {l:.1e}), adjoint {it_adj} "
          f"iterations {t_adj:.2f} s", file=sys.stderr, flush=True)
    th = prob.model.theta()
    fd = {}
    for j in [int(v) for v in args.fd_params.split(",") if v]:
        h = args.fd_step * max(abs(th[j]), 1e-2)
        tp, tm = th.copy(), th.copy()
        tp[j] += h
        tm[j] -= h
        h_eff = float(np.float32(tp[j])) - float(np.float32(tm[j]))
        Lp = solve(prob.model.with_theta(tp))[0]
        Lm = solve(prob.model.with_theta(tm))[0]
        d = (Lp - Lm) / h_eff
        rel = abs(grad[j] - d) / max(abs(d), 1e-300)
        # one-sided differences: where they disagree, L has a kink at theta
        # (an argmin that changes with the sign of the step) and the central
        # difference averages the two sides
        dp = (Lp - L) / (float(np.float32(tp[j])) - float(np.float32(th[j])))
        dm = (L - Lm) / (float(np.float32(th[j])) - float(np.float32(tm[j])))
        fd[j] = {"adjoint": float(grad[j]), "fd": d, "rel": rel, "fd_plus": dp, "fd_minus": dm}
        print(f"# theta[{j}] = {th[j]:+.5g}: adjoint {grad[j]:+.6e} fd {d:+.6e} rel {rel:.1e} "
              f"(one-sided {dp:+.6e} / {dm:+.6e})", file=sys.stderr, flush=True)
    return {"metric": "adjoint dL/dtheta vs central differences", "L": L, "grad": [float(v) for v in grad],
            "f64_sweeps": its, "f64_seconds": t_solve, "adjoint_seconds": t_adj, "adjoint_iterations": it_adj,
            "fd_step_rel": args.fd_step, "fd": fd, "max_rel": max((v["rel"] for v in fd.values()), default=None),
            "config": {"workload": f"problem {args.config}", "grid": list(g.n), "states": g.ncells,
                       "controls": prob.nctl, "starts": 4096}}


def run_ooc(pg, prob, args) -> dict:
    """Tiled sweeps from pinned host memory against in-HBM sweeps."""
    g = prob.grid
    tile = tuple(int(v) for v in args.ooc_tile.split(","))
    t = time.perf_counter()
    vstore = pg.abi.PG_VSTORE_U16_TILE if args.ooc_store == "u16" else pg.abi.PG_VSTORE_F32
    if args.reference_scheme:
        # the reference's default solve.jl scheme, tiled (as run_reference_scheme in HBM)
        from dataclasses import replace
        prob = replace(prob, params=reference_scheme_params(prob.params), ctl=pg.control_set(4))
    s = pg.HipSolver(g, prob.model, prob.params, prob.ctl, prob.occ, vstore=vstore, ooc_budget=int(args.ooc_budget),
                     ooc_tile=tile)
    if args.reference_scheme:
        s.set_warm_start(True)
    info = s.ooc_info()
    print(f"# ooc: tiles {info[0]} x {info[1]}, {info[2] / 1e9:.2f} GB on the device, amplification {info[3]:.2f}, "
          f"set up in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    tl = s.ooc_tiles()
    wl = tl[:, 9:12]
    print(f"# ooc windows: x {wl[:, 0].min()}..{wl[:, 0].max()} (median {int(np.median(wl[:, 0]))}), "
          f"y {wl[:, 1].min()}..{wl[:, 1].max()}, heading {wl[:, 2].min()}..{wl[:, 2].max()} "
          f"(median {int(np.median(wl[:, 2]))})", file=sys.stderr, flush=True)
    big = np.argsort(-wl.prod(axis=1))[:3]
    for k in big:
        print(f"#   tile {k}: interior {tl[k, 0:3].tolist()}+{tl[k, 3:6].tolist()} window {tl[k, 6:9].tolist()}+"
              f"{tl[k, 9:12].tolist()}", file=sys.stderr, flush=True)
    if args.steps == 0:
        s.close()
        return {"ooc_windows_only": True}
    def note(msg):
        print(f"# ooc {msg} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    s.fill_seed(prob.seeds)
    note("seeded")
    ref = None
    if not args.no_ooc_compare:
        ref = pg.HipSolver(g, prob.model, prob.params, prob.ctl, prob.occ, vstore=vstore)
        if args.reference_scheme:
            ref.set_warm_start(True)
        elif vstore == pg.abi.PG_VSTORE_F32:
            ref.set_kernel(pg.abi.PG_KERNEL_LDS)
        ref.set_active(False)
        ref.fill_seed(prob.seeds)
        note("in-HBM reference seeded")
    steps = args.steps if args.steps is not None else 3
    ms, equal = [], []
    for i in range(steps):
        t = time.perf_counter()
        r = s.sweep(i, i % 2 == 1)
        ms.append((time.perf_counter() - t) * 1e3)
        line = f"# ooc sweep {i}: {ms[-1]:.0f} ms, linf {r.linf:.4f}"
        print(line, file=sys.stderr, flush=True)
        if ref is not None:
            rr = ref.sweep(i, i % 2 == 1)
            note(f"in-HBM sweep {i}: {rr.kernel_ms:.0f} ms")
            a = s.download().view(np.uint32)
            note("tiled table downloaded")
            b = ref.download().view(np.uint32)
            note("in-HBM table downloaded")
            same = bool(np.array_equal(a, b))
            del a, b
            equal.append(same and rr.linf == r.linf)
            note(f"sweep {i}: tables {'bit-identical' if equal[-1] else 'DIFFER'}")
    esc_ms, esc_equal = [], []
    if args.ooc_escape > 0:
        if args.ooc_store != "f32":
            s.polish(0)
            if ref is not None:
                ref.polish(0)
        for i in range(args.ooc_escape):
            t = time.perf_counter()
            r = s.escape_sweep(i, i % 2 == 1)
            esc_ms.append((time.perf_counter() - t) * 1e3)
            print(f"# ooc escape sweep {i}: {esc_ms[-1]:.0f} ms, linf {r.linf:.4f}", file=sys.stderr, flush=True)
            if ref is not None:
                rr = ref.escape_sweep(i, i % 2 == 1)
                note(f"in-HBM escape sweep {i}: {rr.kernel_ms:.0f} ms")
                same = bool(np.array_equal(s.download().view(np.uint32), ref.download().view(np.uint32)))
                esc_equal.append(same and rr.linf == r.linf)
                note(f"escape sweep {i}: tables {'bit-identical' if esc_equal[-1] else 'DIFFER'}")
    s.close()
    if ref is not None:
        ref.close()
    extra = {}
    if esc_ms:
        # the first escape sweep also plans the escape windows
        extra["escape"] = {"ms_per_step": esc_ms, "bit_identical_to_in_hbm": (all(esc_equal) if esc_equal else None)}
    return {**extra, "metric": "out-of-core tiled sweep", "store": args.ooc_store, "value": g.ncells / (float(np.mean(ms)) / 1e3),
            "unit": "state-updates/s", "ms_per_step": float(np.mean(ms)), "steps": steps,
            "config": {"workload": f"problem {args.config}", "grid": list(g.n), "states": g.ncells,
                       "controls": prob.nctl, "budget_bytes": int(args.ooc_budget), "tile": list(info[0]),
                       "tiles": info[1], "device_bytes": info[2], "amplification": info[3],
                       "scheme": "reference default (solve.jl)" if args.reference_scheme else "plain"},
            "bit_identical_to_in_hbm": (all(equal) if equal else None)}


def reference_scheme_params(params):
    """The reference's default solve.jl scheme on top of a problem's
    parameters (Run.jl:203-294 defaults)."""
    from dataclasses import replace
    return replace(params, adaptive_checks=True, ntau=5, tau_ratio=2.0, cfl=2.0, tau_min=0.004, tau_max=0.5,
                   ncoarse=8, rounds=2, delta0=0.35, rk2=True, vclamp=False, simplex=True)


def run_reference_scheme(pg, prob, args) -> dict:
    """Sweeps of the reference's default solve.jl scheme on this problem's
    grid, model and obstacles: level-4 octahedron lattice (67 commands) with a
    rotating slice of 8, 2 pattern-search rounds, the Int8 warm start, Kuhn
    lookups, RK2, a per-cell CFL horizon (cfl 2) with a 5-rung ladder and
    adaptive probes (Run.jl:203-294 defaults)."""
    g = prob.grid
    p = reference_scheme_params(prob.params)
    ctl = pg.control_set(4)
    s = pg.HipSolver(g, prob.model, p, ctl, prob.occ)
    s.set_warm_start(True)
    s.fill_seed(prob.seeds)
    for i in range(args.warmup):
        r = s.sweep(i, i % 2 == 1)
        print(f"# reference-scheme warmup sweep {i}: {r.kernel_ms:.1f} ms kernel", file=sys.stderr, flush=True)
    steps = args.steps if args.steps is not None else 5
    kms = []
    t = time.perf_counter()
    for i in range(steps):
        r = s.sweep(args.warmup + i, (args.warmup + i) % 2 == 1)
        kms.append(r.kernel_ms)
        print(f"# reference-scheme sweep {i}: {r.kernel_ms:.1f} ms kernel, linf {r.linf:.4f}", file=sys.stderr,
              flush=True)
    wall = time.perf_counter() - t
    kern, _ = s.kernel_info()
    s.close()
    return {"metric": "cell-updates/s, the reference's default solve.jl scheme", "value": g.ncells * steps / wall,
            "unit": "cell-updates/s", "steps": steps, "warmup": args.warmup, "ms_per_step": wall / steps * 1e3,
            "kernel_ms": float(np.mean(kms)), "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "kernel": {1: "global-gather", 2: "LDS-staged"}.get(kern, str(kern)),
            "config": {"workload": f"problem {args.config} grid, model and obstacles, reference default scheme",
                       "grid": list(g.n), "states": g.ncells, "controls": int(ctl.shape[0]), "control_scan": 8,
                       "refine_rounds": 2, "tau_levels": 5, "cfl": 2.0, "warm_start": True, "simplex": True,
                       "rk2": True, "adaptive_checks": True}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--no-converge", action="store_true",
                    help="skip the time-to-tolerance solves of both forms")
    ap.add_argument("--accuracy", action="store_true",
                    help="also the FP64 table and both forms' subsample errors against it")
    ap.add_argument("--f64-tol", type=float, default=1e-7)
    ap.add_argument("--f64-max-sweeps", type=int, default=1500)
    ap.add_argument("--polish-max", type=int, default=60,
                    help="most exact polish sweeps after the precision form's 16-bit sweeps")
    ap.add_argument("--f64-only", action="store_true")
    ap.add_argument("--adjoint", action="store_true")
    ap.add_argument("--fd-params", default=",".join(str(j) for j in range(30)))
    ap.add_argument("--fd-step", type=float, default=1e-3)
    ap.add_argument("--ooc", action="store_true")
    ap.add_argument("--reference-scheme", action="store_true",
                    help="sweeps of the reference's default solve.jl scheme (global-gather kernel)")
    ap.add_argument("--ooc-budget", type=float, default=8e9)
    ap.add_argument("--ooc-tile", default="16,16,8")
    ap.add_argument("--ooc-store", choices=["f32", "u16"], default="f32",
                    help="the tiled sweep's table: FP32, or the precision form's 16-bit codes")
    ap.add_argument("--no-ooc-compare", action="store_true")
    ap.add_argument("--ooc-escape", type=int, default=0,
                    help="escape sweeps after the tiled value sweeps (the first plans the escape windows)")
    ap.add_argument("--f64-out", default=None)
    ap.add_argument("--f64-sample", default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--max-sweeps", type=int, default=2000)
    ap.add_argument("--no-active", action="store_true",
                    help="converge runs compute every box of every sweep (no active set)")
    ap.add_argument("--kernel", choices=["auto", "global", "lds", "fast"], default="auto",
                    help="sweep kernel of the FP32 run (exact: global, lds; contracted: fast)")
    ap.add_argument("--no-fp16", action="store_true",
                    help="skip the precision-form and contracted-FP32 throughput runs")
    ap.add_argument("--contracted-steps", type=int, default=5,
                    help="timed sweeps of the FP32 store under the contracted kernel")
    ap.add_argument("--escape-steps", type=int, default=3,
                    help="escape sweeps timed after the FP32 throughput steps")
    ap.add_argument("--cpu-converge", action="store_true",
                    help="time the oracle on problem 1 to tolerance, in place and Jacobi, on the host")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a device copy of 1 GiB (256 Mi floats) through torch, the "
                         "reference point that scales FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the solve is single-GPU "
              f"by design and no multi-GPU path is built", file=sys.stderr)
        return 2

    if args.calibrate_copy:
        import torch
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("# copy calibration: 3 x 1 GiB device-to-device copies", file=sys.stderr, flush=True)
        del a, b
        torch.cuda.empty_cache()

    pg = load_package()
    t0 = time.perf_counter()
    prob = pg.configs.problem(args.config)
    g = prob.grid
    print(f"# problem {args.config}: grid {g.n} = {g.ncells} states, {prob.nctl} controls, "
          f"occupied {float(prob.occ.mean()):.3f}, built in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    if args.ooc:
        print(json.dumps(run_ooc(pg, prob, args)), flush=True)
        return 0
    if args.reference_scheme:
        print(json.dumps(run_reference_scheme(pg, prob, args)), flush=True)
        return 0
    if args.adjoint:
        args.f64_tol = min(args.f64_tol, 1e-12)
        print(json.dumps(run_adjoint(pg, prob, args)), flush=True)
        return 0
    if args.f64_only:
        idx, V64, P64, info = f64_sample(pg, prob, args)
        if args.f64_out:
            np.savez(args.f64_out, idx=idx, v64=V64, p64=P64, info=json.dumps(info), config=args.config)
        print(json.dumps({"f64": info, "config": {"workload": f"problem {args.config}", "grid": list(g.n)}}),
              flush=True)
        return 0
    if args.cpu_converge:
        print(json.dumps({"cpu_converge": cpu_converge(pg, args)}), flush=True)
        return 0

    def timed(vstore, kernel, nsteps=None):
        s = pg.HipSolver(g, prob.model, prob.params, prob.ctl, prob.occ, vstore=vstore)
        # throughput is that of sweeps computing every box (an early sweep
        # of a solve would otherwise skip most of the grid)
        s.set_active(False)
        if kernel != "auto":
            s.set_kernel({"global": pg.abi.PG_KERNEL_GLOBAL, "lds": pg.abi.PG_KERNEL_LDS,
                          "fast": pg.abi.PG_KERNEL_FAST}[kernel])
        s.fill_seed(prob.seeds)
        tag = ("fp32" if kernel != "fast" else "fp32_contracted") if vstore == pg.abi.PG_VSTORE_F32 else "u16_tile"
        warm_ms = []
        for i in range(args.warmup):
            r = s.sweep(i, i % 2 == 1)
            warm_ms.append(r.kernel_ms)
            print(f"# {tag} warmup sweep {i}: {r.kernel_ms:.1f} ms kernel", file=sys.stderr, flush=True)
        steps = nsteps if nsteps is not None else args.steps
        if steps is None:
            per = (warm_ms[-1] / 1e3) if warm_ms else 5.0
            steps = int(max(3, min(20, 45.0 / max(per, 1e-3))))
        kms = []
        t = time.perf_counter()
        for i in range(steps):
            r = s.sweep(args.warmup + i, (args.warmup + i) % 2 == 1)
            kms.append(r.kernel_ms)
            print(f"# {tag} sweep {i}: {r.kernel_ms:.1f} ms kernel, linf {r.linf:.4f}", file=sys.stderr,
                  flush=True)
        wall = time.perf_counter() - t
        return s, steps, wall, kms

    s, steps, wall, kms = timed(pg.abi.PG_VSTORE_F32, args.kernel)
    ms_per_step = wall / steps * 1e3
    value = g.ncells * steps / wall
    kavg = float(np.mean(kms))
    kern, coverage = s.kernel_info()
    chords_free, probes_pass = s.probe_info()
    p = prob.params
    flops_update = pg.configs.algorithmic_flops_per_update(prob.nctl, p.substeps(), p.checks, prob.model.knee > 0)
    bytes_update = pg.configs.algorithmic_bytes_per_update(prob.nctl, 4)
    tflops = g.ncells * flops_update / (kavg / 1e3) / 1e12
    pmc = PMC.get((args.config, kern), {})
    traffic = pmc.get("hbm_bytes")
    roofline = {
        "bound": "valu",
        "achieved": tflops,
        "peak": FP32_VECTOR_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": tflops / FP32_VECTOR_PEAK_TFLOPS,
        "traffic": traffic,
        "kernel_ms": kavg,
        "flops_per_state_update": flops_update,
        "flops_basis": "the reference's plain backup counted on its source: every +, -, *, /, sqrt "
                       "(sincos evaluated in double), per state-update",
        "hbm": {"achieved": (traffic / (kavg / 1e3) / 1e9) if traffic else None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": (traffic / (kavg / 1e3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
                "basis": "rocprofv3 FETCH_SIZE + WRITE_SIZE per launch (carried)"},
        "logical_gather": {"bytes_per_state_update": bytes_update,
                           "rate": g.ncells * bytes_update / (kavg / 1e3) / 1e9, "unit": "GB/s",
                           "note": "every corner read counted once, no reuse; served from LDS-staged windows, "
                                   "so this is not HBM traffic and may exceed the HBM peak"},
    }
    if "flops_counted" in pmc:
        roofline["counters"] = {"flops_executed_per_launch": pmc["flops_counted"],
                                "executed_tflops": pmc["flops_counted"] / (kavg / 1e3) / 1e12,
                                **{k: v for k, v in pmc.items() if k not in ("hbm_bytes", "flops_counted")}}
    out = {
        "metric": "state-updates/s + time to ||dV||inf<1e-3 s: FP16-V vs FP32 plain, 1 GPU",
        "value": value,
        "unit": "state-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"problem {args.config}: plain FP32 Jacobi value-iteration sweep",
            "grid": list(g.n),
            "states": g.ncells,
            "controls": prob.nctl,
            "obstacles": len(prob.rects),
            "scheme": f"fixed dt {p.dt:g} s, forward Euler {p.substeps()} substeps, {p.checks} swept probes, "
                      f"64-corner multilinear",
            "parallelism": "single GPU",
            "kernel": {1: "global-gather", 2: "LDS-staged", 3: "LDS-staged, contracted FP32"}.get(kern, str(kern)),
            "window_coverage": coverage,
            "boxes_probe_chords_free": chords_free,
            "boxes_probes_skipped": probes_pass,
        },
        "roofline": roofline,
    }
    if not args.no_cpu:
        V = s.download()
        try:
            out["cpu_baseline"] = cpu_baseline(pg, prob, V)
        except Exception as e:  # the line must still print
            out["cpu_baseline"] = {"error": repr(e)}
    if args.escape_steps > 0:
        # escape sweeps over the table the timed sweeps left: most cells have
        # no route yet, so nearly every cell is backed up (the pass's worst case)
        ems = []
        for i in range(args.escape_steps):
            r = s.escape_sweep(i, i % 2 == 1)
            ems.append(r.kernel_ms)
            print(f"# escape sweep {i}: {r.kernel_ms:.1f} ms kernel, linf {r.linf:.4f}", file=sys.stderr, flush=True)
        out["escape"] = {"kernel_ms": float(np.mean(ems)), "steps": len(ems),
                         "ratio_to_value_sweep": float(np.mean(ems)) / kavg,
                         "kernel": "LDS-staged" if prob.nctl <= 128 else "global-gather"}
    s.close()
    if not args.no_fp16 and prob.nctl <= 128:
        s16, st16, wall16, kms16 = timed(pg.abi.PG_VSTORE_U16_TILE, "auto")
        v16 = g.ncells * st16 / wall16
        out["u16_tile"] = {
            "value": v16, "unit": "state-updates/s", "steps": st16,
            "ms_per_step": wall16 / st16 * 1e3, "kernel_ms": float(np.mean(kms16)),
            "speedup_vs_fp32": v16 / value,
            "store": "16-bit fixed-point codes per 4^6-cell tile with an FP32 (offset, scale) per tile for the "
                     "gathers; each state's own value in FP32; contracted FP32 backups with the diagonal solved",
            "tflops_algorithmic": g.ncells * flops_update / (float(np.mean(kms16)) / 1e3) / 1e12,
        }
        s16.close()
        sc, stc, wallc, kmsc = timed(pg.abi.PG_VSTORE_F32, "fast", nsteps=args.contracted_steps)
        vc = g.ncells * stc / wallc
        out["fp32_contracted"] = {
            "value": vc, "unit": "state-updates/s", "steps": stc, "ms_per_step": wallc / stc * 1e3,
            "kernel_ms": float(np.mean(kmsc)), "speedup_vs_fp32": vc / value,
            "store": "FP32 (Jacobi pair), swept by the precision form's contracted-arithmetic kernel",
        }
        out["u16_tile"]["speedup_vs_fp32_contracted"] = v16 / vc
        sc.close()
    if not args.no_converge or args.accuracy:
        try:
            out["converge"], tables = converge_both(pg, prob, args)
            if args.accuracy:
                out["accuracy"] = accuracy(pg, prob, args, tables)
        except Exception as e:  # the line must still print
            out["converge_error"] = repr(e)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
